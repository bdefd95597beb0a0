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
//  * The mirror's upkeep runs on a background thread of its own, in the
//    order the device applied the key sets, overlapping the device's index
//    maintenance, masks and its share of the DFS; the host count waits for
//    it before reading the mirror.
//  * The share has no floor: the reference clamps it to [0.05, 0.95], which
//    on a host that counts hundreds of times slower than the GPU would hand
//    it 5 % of the rows, and the batch 30 times the GPU's time, every batch.
//    Here it stays below 0.95, and the balance it moves toward also charges
//    the host's wait for the mirror: with rates c (host) and g (GPU) in rows
//    per ms and a wait of w ms, the host's k rows end with the GPU's n - k
//    when w + k / c = (n - k) / g.  A share under one row leaves the batch
//    to the GPU; after 16 such batches one row per host thread probes the
//    rates again.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <functional>
#include <memory>
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
    uint32_t idle = 0;  // batches left to the GPU since the last split
    // Host mirror of M: row = owner * ndir + dir -> sorted targets.  The
    // flags say what the queued upkeep will have built; reading the mirror
    // waits for the queue (Drain / WaitIdle).
    bool m_valid = false;
    std::vector<std::vector<uint32_t>> m;
    std::atomic<uint64_t> m_entries{0};
    // U of the current batch: sorted targets with duplicates, and each
    // touched row's range in them.
    bool u_valid = false;
    std::vector<uint32_t> u_vals;
    std::unordered_map<uint64_t, std::pair<uint64_t, uint32_t>> u_rows;
    std::vector<uint64_t> upd_keys;  // the batch's Upd keys, sorted (committed into M)
    // Background upkeep: one thread running queued tasks in order.
    std::thread keeper;
    std::mutex kmu;
    std::condition_variable kcv;
    std::deque<std::function<void()>> tasks;
    bool kbusy = false, kquit = false;
    std::exception_ptr kerr;  // first failed task, reported by the next Drain
    std::atomic<double> mirror_ms{0};  // upkeep time of the batch's tasks so far
    // Host copies for the count.
    std::vector<uint64_t> masks;
    std::vector<uint32_t> rows;
    // Running host count.
    std::vector<std::thread> workers;
    std::atomic<uint64_t> next{0};
    std::atomic<uint32_t> stop{0};  // 1: abort asked, 2: deadline passed, 3: mirror upkeep failed
    std::atomic<uint32_t> bad{0};   // a row held an out-of-range tag or vertex id
    std::mutex mu;                  // guards per_edge, wide and the times below
    uint64_t per_edge[kMaxQE] = {};
    bool wide = false;
    std::chrono::steady_clock::time_point t0, t_count, t_end;
    bool counting = false;
    uint64_t k = 0, n = 0;
    uint32_t D = 0;
    // Statistics.
    uint64_t batches = 0, last_cpu_rows = 0, last_gpu_rows = 0, last_cpu_matches = 0;
    double last_cpu_ms = 0, last_gpu_ms = 0, last_wait_ms = 0;
};

}  // namespace cgd

namespace cgapi {

namespace {

using Clock = std::chrono::steady_clock;

double MsSince(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }
double MsBetween(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

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

// ---------------------------------------------------------------- upkeep queue
void KeeperLoop(SplitState* S) {
    std::unique_lock<std::mutex> lk(S->kmu);
    for (;;) {
        S->kcv.wait(lk, [&] { return S->kquit || !S->tasks.empty(); });
        if (S->tasks.empty()) return;  // quit once the queue is empty
        std::function<void()> t = std::move(S->tasks.front());
        S->tasks.pop_front();
        S->kbusy = true;
        lk.unlock();
        try {
            t();
        } catch (...) {
            lk.lock();
            if (!S->kerr) S->kerr = std::current_exception();
            lk.unlock();
        }
        lk.lock();
        S->kbusy = false;
        S->kcv.notify_all();
    }
}

void Post(SplitState& S, std::function<void()> t) {
    std::lock_guard<std::mutex> g(S.kmu);
    if (!S.keeper.joinable()) S.keeper = std::thread(KeeperLoop, &S);
    S.tasks.push_back(std::move(t));
    S.kcv.notify_all();
}

// Waits until every queued task ran; false if one of them failed.
bool WaitIdle(SplitState& S) {
    std::unique_lock<std::mutex> lk(S.kmu);
    S.kcv.wait(lk, [&] { return S.tasks.empty() && !S.kbusy; });
    return !S.kerr;
}

// WaitIdle, rethrowing a failed task's error (the mirror is then dropped).
void Drain(SplitState& S) {
    if (WaitIdle(S)) return;
    std::exception_ptr e;
    {
        std::lock_guard<std::mutex> g(S.kmu);
        std::swap(e, S.kerr);
    }
    S.m_valid = S.u_valid = false;
    std::rethrow_exception(e);
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

// M <- M | keys (remove = false) or M \ keys (remove = true).
void MirrorMerge(const cgbdsm_query* q, SplitState& S, std::vector<uint64_t>& keys, bool remove) {
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
    int64_t sum = 0;
    for (int64_t d : delta) sum += d;
    S.m_entries.fetch_add((uint64_t)sum);
}

// Mirror of M built from the device's export (the index after the last
// commit), on the calling thread once the queue is idle.
void MirrorBuild(cgbdsm_query* q) {
    SplitState& S = *q->split;
    Drain(S);
    uint64_t n = 0;
    if (cgbdsm_index_export(q, 0, nullptr, 0, &n) != CGBDSM_OK) Throw(CGBDSM_ERR_DEVICE, g_err);
    std::vector<uint64_t> keys(n);
    if (n && cgbdsm_index_export(q, 0, keys.data(), n, &n) != CGBDSM_OK) Throw(CGBDSM_ERR_DEVICE, g_err);
    S.m.clear();
    S.m_entries.store(0);
    MirrorMerge(q, S, keys, false);
    S.m_valid = true;
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
    // The mirror must hold this batch's M and U before any row is read.
    if (!WaitIdle(S)) {
        uint32_t z = 0;
        S.stop.compare_exchange_strong(z, 3);
    }
    {
        std::lock_guard<std::mutex> g(S.mu);
        if (!S.counting) {
            S.counting = true;
            S.t_count = Clock::now();
        }
    }
    HostDfs H{q, &S};
    uint64_t local[kMaxQE] = {};
    const uint32_t D = S.D, ts = q->hplan.tag_shift, vm = q->hplan.v_mask;
    for (;;) {
        if (H.Stopped(&S)) break;
        const uint64_t r = S.next.fetch_add(1, std::memory_order_relaxed);
        if (r >= S.k) break;
        const uint32_t* row = S.rows.data() + r * D;
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
    const auto now = Clock::now();
    if (now > S.t_end) S.t_end = now;
}

}  // namespace

void SplitJoin(cgbdsm_query* q, bool abort) {
    SplitState& S = *q->split;
    if (abort) {
        uint32_t z = 0;
        S.stop.compare_exchange_strong(z, 1);
    }
    for (auto& w : S.workers) w.join();
    S.workers.clear();
}

void SplitFree(cgbdsm_query* q) {
    if (!q->split) return;
    SplitState& S = *q->split;
    SplitJoin(q, true);
    {
        std::lock_guard<std::mutex> g(S.kmu);
        S.kquit = true;
        S.kcv.notify_all();
    }
    if (S.keeper.joinable()) S.keeper.join();
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
    S.u_valid = false;
    if (!S.m_valid) MirrorBuild(q);
    auto ins = std::make_shared<std::vector<uint64_t>>();
    auto del = std::make_shared<std::vector<uint64_t>>();
    CopyKeys(q, q->k_ins, q->n_ins, *ins);
    CopyKeys(q, q->k_del, q->n_del, *del);
    Post(S, [q, ins, del] {
        SplitState& St = *q->split;
        const auto t0 = Clock::now();
        MirrorMerge(q, St, *ins, false);
        MirrorMerge(q, St, *del, true);
        St.mirror_ms.store(MsSince(t0));
    });
}

void SplitOnUpdateIndex(cgbdsm_query* q) {
    SplitState& S = *q->split;
    auto upd = std::make_shared<std::vector<uint64_t>>();
    CopyKeys(q, q->k_upd, q->n_upd, *upd);
    Post(S, [q, upd] {
        SplitState& St = *q->split;
        const auto t0 = Clock::now();
        std::vector<uint64_t>& K = *upd;
        std::sort(K.begin(), K.end());
        St.u_vals.resize(K.size());
        St.u_rows.clear();
        const std::vector<uint64_t> starts = Groups(K);
        for (uint64_t g = 0; g + 1 < starts.size(); g++)
            St.u_rows[RowOf(q, K[starts[g]])] = {starts[g], (uint32_t)(starts[g + 1] - starts[g])};
        for (uint64_t i = 0; i < K.size(); i++) St.u_vals[i] = (uint32_t)K[i];
        St.upd_keys.swap(K);
        St.mirror_ms.store(St.mirror_ms.load() + MsSince(t0));
    });
    S.u_valid = S.m_valid;
}

void SplitOnCommit(cgbdsm_query* q) {
    SplitState& S = *q->split;
    // M takes the Upd keys the update-index task saved (it ran before).
    if (S.m_valid && S.u_valid) {
        Post(S, [q] {
            SplitState& St = *q->split;
            const auto t0 = Clock::now();
            MirrorMerge(q, St, St.upd_keys, false);
            St.mirror_ms.store(St.mirror_ms.load() + MsSince(t0));
        });
    } else {
        S.m_valid = false;
    }
    S.u_valid = false;
}

uint64_t SplitPlan(cgbdsm_query* q, uint64_t n) {
    SplitState& S = *q->split;
    if (!S.threads || !S.m_valid || !S.u_valid || n < std::max<uint64_t>(S.min_rows, 2)) return 0;
    uint64_t k = S.measured || S.ratio0 > 0 ? (uint64_t)(S.ratio * (double)n) : (uint64_t)S.threads;
    if (!k && ++S.idle >= 16) k = (uint64_t)S.threads;  // probe the rates again
    if (k) S.idle = 0;
    return std::min(k, n - 1);
}

void SplitStart(cgbdsm_query* q, const uint32_t* d_rows, uint64_t k, uint64_t n, uint32_t D) {
    SplitState& S = *q->split;
    S.rows.resize(k * D);
    CGD_HIP(hipMemcpyAsync(S.rows.data(), d_rows, k * D * 4, hipMemcpyDeviceToHost, q->s));
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
    S.counting = false;
    S.t0 = S.t_count = S.t_end = Clock::now();
    const int t = (int)std::min<uint64_t>((uint64_t)S.threads, k);
    for (int i = 0; i < t; i++) S.workers.emplace_back(HostWorker, q);
}

void SplitFinish(cgbdsm_query* q, double gpu_ms) {
    SplitState& S = *q->split;
    SplitJoin(q, false);
    Drain(S);
    if (S.bad.load())
        Throw(CGBDSM_ERR_DEVICE, "host share of the final DFS read a frontier row with an out-of-range tag or vertex id");
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
    const double wait_ms = MsBetween(S.t0, S.t_count), cpu_ms = MsBetween(S.t_count, S.t_end);
    const uint64_t gpu_rows = S.n - S.k;
    if (S.alpha > 0 && cpu_ms > 1e-6 && gpu_ms > 1e-6 && gpu_rows) {
        // Rows per ms of each side; the host's k rows end with the GPU's
        // n - k when wait + k / c = (n - k) / g.
        const double c = (double)S.k / cpu_ms, g = (double)gpu_rows / gpu_ms, nn = (double)S.n;
        const double kb = std::max(0.0, (nn / g - wait_ms) / (1.0 / c + 1.0 / g));
        const double rt = kb / nn;
        S.ratio = S.measured || S.ratio0 > 0 ? S.alpha * rt + (1.0 - S.alpha) * S.ratio : rt;
        S.ratio = std::min(S.ratio, 0.95);
        S.measured = true;
    }
    S.batches++;
    S.last_cpu_rows = S.k;
    S.last_gpu_rows = gpu_rows;
    S.last_cpu_ms = cpu_ms;
    S.last_wait_ms = wait_ms;
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
        Drain(S);
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
        out->last_wait_ms = S.last_wait_ms;
        out->last_cpu_matches = S.last_cpu_matches;
        out->mirror_entries = S.m_valid ? S.m_entries.load() : 0;
        out->mirror_ms = S.mirror_ms.load();
    });
}

cgbdsm_status cgbdsm_cpu_mirror_export(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out) {
    return Guard([&] {
        if (!q || (which != 0 && which != 1)) Throw(CGBDSM_ERR_INVALID, "bad query or index selector");
        if (!q->split) Throw(CGBDSM_ERR_INVALID, "the CPU split is off (cgbdsm_set_cpu_split)");
        SplitState& S = *q->split;
        Drain(S);
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
