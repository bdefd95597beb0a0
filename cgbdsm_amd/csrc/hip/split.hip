// CPU/GPU split of the final DFS: the host counts the first k rows of a
// query's frontier on its own cores while the GPU counts rows [k, n).
//
// Reference behaviour (graph/graph_gpu.cu:1185-1258, Step5CPUGPUSplit;
// index/cpu_dfs.cpp:173-217, cpuExtendBFSAllBitCountParallel; the host index
// mirror index/cpu_index_mirror.cpp, kept in sync at
// index/calig_helper_gamma.cu:279-302, 503-516, 558-566, 588-595): rows
// [0, k) go to the host, k = ratio * n, nothing is split below a minimum
// frontier, and after each split batch the ratio moves toward the measured
// balance cpu_rate / (cpu_rate + gpu_rate) by an exponential moving average.
//
// Here:
//  * The host mirror holds the main index M and the batch's update index U
//    as sorted lists per (owner, query direction).  It is built once from
//    the device's own export of M, then follows the key sets the device
//    applies: at transfer_update_all M <- (M | Ins) \ Del, at
//    transfer_update_index U <- Upd (duplicates kept, as the device's U
//    keeps them), at the commit M <- M | Upd.  A (re)initialised index
//    marks it stale; it is rebuilt at the next batch.
//  * The host count follows the device's rules exactly (support-mask bit,
//    injectivity, membership in every backward neighbour's list, each list
//    read with the edge tag's visibility: U for the tag's own two
//    directions, M plus appended U for directions of earlier edges), so its
//    per-edge counts add to the GPU's.
//  * The ratio has no floor: the reference clamps it to [0.05, 0.95], which
//    on a host that counts hundreds of times slower than the GPU would hand
//    it 5 % of the rows, and the batch 30 times the GPU's time, every batch.
//    Here it only stays below 0.95; a share under one row leaves the batch
//    to the GPU.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <dirent.h>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "api.h"

namespace cgd {

struct SplitState {
    int threads = 0;
    double ratio = 0, ratio0 = 0, alpha = 0.3;
    uint64_t min_rows = 8192;
    bool measured = false;
    // Host mirror of M: row = owner * ndir + dir -> sorted targets.
    bool m_valid = false;
    std::vector<std::vector<uint32_t>> m;
    uint64_t m_entries = 0;
    // U of the current batch: sorted (row, target) pairs with duplicates,
    // and each touched row's range in them.
    bool u_valid = false;
    std::vector<uint32_t> u_vals;
    std::unordered_map<uint64_t, std::pair<uint32_t, uint32_t>> u_rows;
    // Host copies for the count.
    std::vector<uint64_t> masks;
    uint32_t* h_rows = nullptr;  // pinned
    uint64_t h_rows_cap = 0;     // u32 words
    std::vector<uint64_t> keys;      // key staging
    std::vector<uint64_t> upd_keys;  // the batch's Upd keys, sorted (committed into M)
    // Running host count.
    std::vector<std::thread> workers;
    std::atomic<uint64_t> next{0};
    std::atomic<uint32_t> stop{0};  // 1: abort asked, 2: deadline passed
    std::atomic<uint32_t> bad{0};   // a row held an out-of-range tag or vertex id
    std::mutex mu;                  // guards per_edge and wide
    uint64_t per_edge[kMaxQE] = {};
    bool wide = false;
    std::chrono::steady_clock::time_point t0;
    double cpu_ms = 0;
    uint64_t k = 0, n = 0;
    uint32_t D = 0;
    // Statistics.
    uint64_t batches = 0, last_cpu_rows = 0, last_gpu_rows = 0, last_cpu_matches = 0;
    double last_cpu_ms = 0, last_gpu_ms = 0, mirror_ms = 0;
};

}  // namespace cgd

namespace cgapi {

namespace {

// Threads of the reference's host rule: hardware threads / NUMA nodes
// (index/calig.cpp:1092-1095).
int DefaultThreads() {
    int nodes = 0;
    if (DIR* d = opendir("/sys/devices/system/node")) {
        while (dirent* e = readdir(d))
            if (!strncmp(e->d_name, "node", 4) && e->d_name[4] >= '0' && e->d_name[4] <= '9') nodes++;
        closedir(d);
    }
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    return std::max(1, hw / std::max(1, nodes));
}

// Runs f(i) for i in [0, n) on up to `threads` threads (contiguous chunks).
template <class F>
void ParallelFor(uint64_t n, int threads, F&& f) {
    const uint64_t t = std::min<uint64_t>((uint64_t)std::max(1, threads), std::max<uint64_t>(n / 256, 1));
    if (t <= 1) {
        for (uint64_t i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (uint64_t w = 0; w < t; w++)
        th.emplace_back([&, w] {
            for (uint64_t i = n * w / t; i < n * (w + 1) / t; i++) f(i);
        });
    for (auto& x : th) x.join();
}

void CopyKeys(cgbdsm_query* q, const DBuf<uint64_t>& b, uint64_t n, std::vector<uint64_t>& out) {
    out.resize(n);
    if (n) {
        CGD_HIP(hipMemcpyAsync(out.data(), b.p, n * 8, hipMemcpyDeviceToHost, q->s));
        CGD_HIP(hipStreamSynchronize(q->s));
    }
}

uint64_t RowOf(const cgbdsm_query* q, uint64_t key) {
    const uint64_t e = key >> q->ushift, u = (key >> 32) & q->umask;
    return u * q->ndir + e;
}

// Groups of equal (dir, owner) in sorted keys: [starts[i], starts[i+1]).
std::vector<uint64_t> Groups(const std::vector<uint64_t>& keys) {
    std::vector<uint64_t> starts;
    for (uint64_t i = 0; i < keys.size(); i++)
        if (!i || (keys[i] >> 32) != (keys[i - 1] >> 32)) starts.push_back(i);
    starts.push_back(keys.size());
    return starts;
}

// M <- M | keys (remove = false) or M \ keys (remove = true), keys sorted.
void MirrorMerge(cgbdsm_query* q, std::vector<uint64_t>& keys, bool remove) {
    SplitState& S = *q->split;
    if (keys.empty()) return;
    std::sort(keys.begin(), keys.end());
    const std::vector<uint64_t> starts = Groups(keys);
    const uint64_t ng = starts.size() - 1;
    uint64_t top = 0;
    for (uint64_t g = 0; g < ng; g++) top = std::max(top, RowOf(q, keys[starts[g]]) + 1);
    if (top > S.m.size()) S.m.resize(top);
    std::vector<int64_t> delta(ng, 0);
    ParallelFor(ng, S.threads, [&](uint64_t g) {
        std::vector<uint32_t>& L = S.m[RowOf(q, keys[starts[g]])];
        std::vector<uint32_t> v;
        for (uint64_t i = starts[g]; i < starts[g + 1]; i++)
            if (v.empty() || v.back() != (uint32_t)keys[i]) v.push_back((uint32_t)keys[i]);
        std::vector<uint32_t> out;
        out.reserve(remove ? L.size() : L.size() + v.size());
        if (remove) std::set_difference(L.begin(), L.end(), v.begin(), v.end(), std::back_inserter(out));
        else std::set_union(L.begin(), L.end(), v.begin(), v.end(), std::back_inserter(out));
        delta[g] = (int64_t)out.size() - (int64_t)L.size();
        L.swap(out);
    });
    for (int64_t d : delta) S.m_entries += d;
}

// Mirror of M built from the device's export (the index after the last commit).
void MirrorBuild(cgbdsm_query* q) {
    SplitState& S = *q->split;
    uint64_t n = 0;
    if (cgbdsm_index_export(q, 0, nullptr, 0, &n) != CGBDSM_OK) Throw(CGBDSM_ERR_DEVICE, g_err);
    S.keys.assign(n, 0);
    if (n && cgbdsm_index_export(q, 0, S.keys.data(), n, &n) != CGBDSM_OK) Throw(CGBDSM_ERR_DEVICE, g_err);
    S.m.clear();
    S.m_entries = 0;
    MirrorMerge(q, S.keys, false);
    S.m_valid = true;
}

double MsSince(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

// ---------------------------------------------------------------- host count
struct HList {
    const uint32_t* base = nullptr;
    const uint32_t* app = nullptr;
    uint32_t bl = 0, al = 0;
};

bool Has(const uint32_t* a, uint32_t n, uint32_t c) { return std::binary_search(a, a + n, c); }

struct HostDfs {
    const cgbdsm_query* q;
    const SplitState* S;
    uint32_t ei = 0, sa = 0, sb = 0;
    uint32_t matched[kMaxQV] = {};
    bool wide = false;
    uint64_t polls = 0;

    HList M(uint64_t row) const {
        HList L;
        if (row < S->m.size()) {
            L.base = S->m[row].data();
            L.bl = (uint32_t)S->m[row].size();
        }
        return L;
    }
    void U(uint64_t row, const uint32_t*& p, uint32_t& n) const {
        auto it = S->u_rows.find(row);
        if (it == S->u_rows.end()) {
            p = nullptr;
            n = 0;
        } else {
            p = S->u_vals.data() + it->second.first;
            n = it->second.second;
        }
    }
    // List of query direction `dir` at data vertex `owner` under tag ei.
    HList Resolve(uint32_t dir, uint32_t owner) const {
        const uint64_t row = (uint64_t)owner * q->ndir + dir;
        HList L;
        if (dir == sa || dir == sb) U(row, L.base, L.bl);
        else L = M(row);
        if (q->plan.dir_to_edge[dir] < ei) U(row, L.app, L.al);
        return L;
    }
    bool Tile(uint32_t qv, uint32_t v) const {
        if (q->hplan.all_ones[ei][qv]) return true;
        return (S->masks[((uint64_t)ei * kMaxQV + qv) * q->W + (v >> 6)] >> (v & 63)) & 1ull;
    }
    // Stop asked, or the deadline passed (then asks every worker to stop).
    bool Stopped(SplitState* St) const {
        if (St->stop.load(std::memory_order_relaxed)) return true;
        if (q->deadline_ms > 0 && MsSince(St->t0) > q->deadline_ms) {
            uint32_t z = 0;
            St->stop.compare_exchange_strong(z, 2);
            return true;
        }
        return false;
    }
    static uint64_t Add(uint64_t a, uint64_t b, bool& w) {
        const uint64_t s = a + b;
        if (s < a) {
            w = true;
            return ~0ull;
        }
        return s;
    }
    // Embeddings extending positions 0..d-1 of the global order.
    uint64_t Extend(uint32_t d, SplitState* St) {
        const cgbdsm_plan& p = q->plan;
        const uint32_t ud = p.go_vs[d];
        const uint32_t s = p.go_bni_offs[d], e = p.go_bni_offs[d + 1];
        HList L[kMaxQV];
        uint32_t nl = 0, scan = 0;
        for (uint32_t k = s; k < e; k++, nl++) {
            const uint32_t bq = p.go_bni[k];
            L[nl] = Resolve(p.eidx[p.go_vs[bq] * kMaxQV + ud], matched[bq]);
            if (L[nl].bl + L[nl].al < L[scan].bl + L[scan].al) scan = nl;
        }
        uint64_t cnt = 0;
        const HList& Ls = L[scan];
        for (uint32_t t = 0; t < Ls.bl + Ls.al; t++) {
            if ((++polls & 0xFFFFu) == 0 && Stopped(St)) return cnt;
            const uint32_t c = t < Ls.bl ? Ls.base[t] : Ls.app[t - Ls.bl];
            if (!Tile(ud, c)) continue;
            bool ok = true;
            for (uint32_t i = 0; i < d && ok; i++) ok = matched[i] != c;
            for (uint32_t k = 0; k < nl && ok; k++)
                if (k != scan) ok = Has(L[k].base, L[k].bl, c) || (L[k].al && Has(L[k].app, L[k].al, c));
            if (!ok) continue;
            if (d + 1 == p.qv) {
                cnt = Add(cnt, 1, wide);
            } else {
                matched[d] = c;
                cnt = Add(cnt, Extend(d + 1, St), wide);
            }
        }
        return cnt;
    }
};

void HostWorker(cgbdsm_query* q) {
    SplitState& S = *q->split;
    HostDfs H{q, &S};
    uint64_t local[kMaxQE] = {};
    const uint32_t D = S.D, ts = q->hplan.tag_shift, vm = q->hplan.v_mask;
    for (;;) {
        if (H.Stopped(&S)) break;
        const uint64_t r = S.next.fetch_add(1, std::memory_order_relaxed);
        if (r >= S.k) break;
        const uint32_t* row = S.h_rows + r * D;
        const uint32_t ei = row[0] >> ts;
        bool good = ei < q->qe;
        H.matched[0] = row[0] & vm;
        for (uint32_t i = 1; i < D; i++) H.matched[i] = row[i];
        for (uint32_t i = 0; i < D && good; i++) good = H.matched[i] < q->dv;
        if (!good) {
            S.bad.store(1);
            S.stop.store(1);
            break;
        }
        H.ei = ei;
        H.sa = q->hplan.sa[ei];
        H.sb = q->hplan.sb[ei];
        const uint64_t c = D >= q->qv ? 1 : H.Extend(D, &S);
        local[ei] = HostDfs::Add(local[ei], c, H.wide);
    }
    std::lock_guard<std::mutex> g(S.mu);
    for (uint32_t e = 0; e < q->qe; e++) S.per_edge[e] = HostDfs::Add(S.per_edge[e], local[e], S.wide);
    S.wide |= H.wide;
}

}  // namespace

void SplitFree(cgbdsm_query* q) {
    if (!q->split) return;
    SplitJoin(q, true);
    if (q->split->h_rows) (void)hipHostFree(q->split->h_rows);
    delete q->split;
    q->split = nullptr;
}

void SplitInvalidate(cgbdsm_query* q) {
    if (!q->split) return;
    q->split->m_valid = false;
    q->split->u_valid = false;
}

void SplitOnApply(cgbdsm_query* q) {
    SplitState& S = *q->split;
    const auto t0 = std::chrono::steady_clock::now();
    S.u_valid = false;
    if (!S.m_valid) MirrorBuild(q);
    CopyKeys(q, q->k_ins, q->n_ins, S.keys);
    MirrorMerge(q, S.keys, false);
    CopyKeys(q, q->k_del, q->n_del, S.keys);
    MirrorMerge(q, S.keys, true);
    S.mirror_ms = MsSince(t0);
}

void SplitOnUpdateIndex(cgbdsm_query* q) {
    SplitState& S = *q->split;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<uint64_t>& K = S.upd_keys;
    CopyKeys(q, q->k_upd, q->n_upd, K);
    std::sort(K.begin(), K.end());
    S.u_vals.resize(K.size());
    S.u_rows.clear();
    const std::vector<uint64_t> starts = Groups(K);
    for (uint64_t g = 0; g + 1 < starts.size(); g++)
        S.u_rows[RowOf(q, K[starts[g]])] = {(uint32_t)starts[g], (uint32_t)(starts[g + 1] - starts[g])};
    for (uint64_t i = 0; i < K.size(); i++) S.u_vals[i] = (uint32_t)K[i];
    S.u_valid = S.m_valid;
    S.mirror_ms += MsSince(t0);
}

void SplitOnCommit(cgbdsm_query* q) {
    SplitState& S = *q->split;
    const auto t0 = std::chrono::steady_clock::now();
    // The Upd keys saved by SplitOnUpdateIndex (commit follows it).
    if (S.m_valid && S.u_valid) MirrorMerge(q, S.upd_keys, false);
    else S.m_valid = false;
    S.u_valid = false;
    S.mirror_ms += MsSince(t0);
}

uint64_t SplitPlan(cgbdsm_query* q, uint64_t n) {
    SplitState& S = *q->split;
    if (!S.threads || !S.m_valid || !S.u_valid || n < std::max<uint64_t>(S.min_rows, 2)) return 0;
    uint64_t k = S.measured || S.ratio0 > 0 ? (uint64_t)(S.ratio * (double)n) : (uint64_t)S.threads;
    return std::min(k, n - 1);
}

void SplitStart(cgbdsm_query* q, const uint32_t* d_rows, uint64_t k, uint64_t n, uint32_t D) {
    SplitState& S = *q->split;
    const uint64_t words = k * D;
    if (words > S.h_rows_cap) {
        if (S.h_rows) CGD_HIP(hipHostFree(S.h_rows));
        S.h_rows = nullptr;
        S.h_rows_cap = 0;
        const uint64_t cap = std::max<uint64_t>(words + words / 2, 1024);
        CGD_HIP(hipHostMalloc((void**)&S.h_rows, cap * 4, hipHostMallocDefault));
        S.h_rows_cap = cap;
    }
    CGD_HIP(hipMemcpyAsync(S.h_rows, d_rows, words * 4, hipMemcpyDeviceToHost, q->s));
    S.masks.resize((uint64_t)q->qe * kMaxQV * q->W);
    if (!S.masks.empty())
        CGD_HIP(hipMemcpyAsync(S.masks.data(), q->masks.p, S.masks.size() * 8, hipMemcpyDeviceToHost, q->s));
    CGD_HIP(hipStreamSynchronize(q->s));
    S.k = k;
    S.n = n;
    S.D = D;
    S.next.store(0);
    S.stop.store(0);
    S.bad.store(0);
    memset(S.per_edge, 0, sizeof(S.per_edge));
    S.wide = false;
    S.t0 = std::chrono::steady_clock::now();
    const int t = (int)std::min<uint64_t>((uint64_t)S.threads, k);
    for (int i = 0; i < t; i++) S.workers.emplace_back(HostWorker, q);
}

void SplitJoin(cgbdsm_query* q, bool abort) {
    SplitState& S = *q->split;
    if (abort) {
        uint32_t z = 0;
        S.stop.compare_exchange_strong(z, 1);
    }
    for (auto& w : S.workers) w.join();
    S.workers.clear();
    S.cpu_ms = MsSince(S.t0);
}

void SplitFinish(cgbdsm_query* q, double gpu_ms) {
    SplitState& S = *q->split;
    SplitJoin(q, false);
    if (S.bad.load()) Throw(CGBDSM_ERR_DEVICE, "host share of the final DFS read a frontier row with an out-of-range tag or vertex id");
    if (S.stop.load() == 2)
        Throw(CGBDSM_ERR_OVERFLOW, "host share of the final DFS exceeded the deadline of " +
                                       std::to_string(q->deadline_ms) + " ms; count incomplete");
    uint64_t total = 0;
    bool wide = S.wide;
    for (uint32_t e = 0; e < q->qe; e++) {
        unsigned long long& pe = q->h_cnt->per_edge[e];
        if (pe + S.per_edge[e] < pe) wide = true;
        pe += S.per_edge[e];
        total += S.per_edge[e];
    }
    if (wide) q->h_cnt->count_ovf = 1;
    const uint64_t gpu_rows = S.n - S.k;
    if (S.alpha > 0 && S.cpu_ms > 1e-6 && gpu_ms > 1e-6 && gpu_rows) {
        const double cpu_tp = (double)S.k / S.cpu_ms, gpu_tp = (double)gpu_rows / gpu_ms;
        const double rt = cpu_tp / (cpu_tp + gpu_tp);
        S.ratio = S.measured || S.ratio0 > 0 ? S.alpha * rt + (1.0 - S.alpha) * S.ratio : rt;
        S.ratio = std::min(S.ratio, 0.95);
        S.measured = true;
    }
    S.batches++;
    S.last_cpu_rows = S.k;
    S.last_gpu_rows = gpu_rows;
    S.last_cpu_ms = S.cpu_ms;
    S.last_gpu_ms = gpu_ms;
    S.last_cpu_matches = total;
}

}  // namespace cgapi

using namespace cgapi;

extern "C" {

cgbdsm_status cgbdsm_set_cpu_split(cgbdsm_query* q, int threads, double ratio, double alpha, uint64_t min_rows) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        if (!(ratio >= 0 && ratio <= 0.95) || !(alpha >= 0 && alpha <= 1))
            Throw(CGBDSM_ERR_INVALID, "ratio must be in [0, 0.95] and alpha in [0, 1]");
        if (threads == 0) {
            SplitFree(q);
            return;
        }
        if (!q->split) q->split = new SplitState();
        SplitState& S = *q->split;
        S.threads = threads < 0 ? DefaultThreads() : threads;
        S.ratio = S.ratio0 = ratio;
        S.alpha = alpha;
        S.min_rows = min_rows ? min_rows : 8192;
        S.measured = false;
    });
}

cgbdsm_status cgbdsm_cpu_split_stats(cgbdsm_query* q, cgbdsm_split_stats* out) {
    return Guard([&] {
        if (!q || !out) Throw(CGBDSM_ERR_INVALID, "null argument");
        memset(out, 0, sizeof(*out));
        if (!q->split) return;
        const SplitState& S = *q->split;
        out->threads = S.threads;
        out->ratio = S.ratio;
        out->batches = S.batches;
        out->last_cpu_rows = S.last_cpu_rows;
        out->last_gpu_rows = S.last_gpu_rows;
        out->last_cpu_ms = S.last_cpu_ms;
        out->last_gpu_ms = S.last_gpu_ms;
        out->last_cpu_matches = S.last_cpu_matches;
        out->mirror_entries = S.m_valid ? S.m_entries : 0;
        out->mirror_ms = S.mirror_ms;
    });
}

cgbdsm_status cgbdsm_cpu_mirror_export(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out) {
    return Guard([&] {
        if (!q || (which != 0 && which != 1)) Throw(CGBDSM_ERR_INVALID, "bad query or index selector");
        if (!q->split) Throw(CGBDSM_ERR_INVALID, "the CPU split is off (cgbdsm_set_cpu_split)");
        const SplitState& S = *q->split;
        if (which == 0 ? !S.m_valid : !S.u_valid) Throw(CGBDSM_ERR_INVALID, "the host mirror holds no such index yet");
        std::vector<uint64_t> h;
        auto emit = [&](uint64_t row, const uint32_t* v, uint64_t n) {
            const uint64_t u = row / q->ndir, e = row % q->ndir;
            for (uint64_t i = 0; i < n; i++) h.push_back((e << q->ushift) | (u << 32) | v[i]);
        };
        if (which == 0) {
            for (uint64_t r = 0; r < S.m.size(); r++) emit(r, S.m[r].data(), S.m[r].size());
        } else {
            for (const auto& kv : S.u_rows) emit(kv.first, S.u_vals.data() + kv.second.first, kv.second.second);
        }
        std::sort(h.begin(), h.end());
        if (n_out) *n_out = h.size();
        if (out) memcpy(out, h.data(), std::min<uint64_t>(cap, h.size()) * 8);
    });
}

}  // extern "C"
