// C ABI of the Stage-2 engine (include/cgbdsm.h).  Each entry point names the
// reference member function it replaces; all device work of a query runs on
// that query's own HIP stream, and no entry point exits the process.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstring>
#include <mutex>
#include <deque>
#include <new>

#include "api.h"

using namespace cgd;
using namespace cgapi;

// Device scratch of cgbdsm_index_export_static, the export entry point that
// allocates nothing: row lengths, the scan's partial sums and total, and the
// exported keys all live in this array (32 MiB on every device the library
// runs on, held until the process ends).  Host code never passes `scratch`
// itself to a kernel: cgbdsm_ctx_create looks up its device address once per
// device (StaticScratch) and every use goes through that pointer.
constexpr uint64_t kStaticScratchWords = 4ull << 20;
static __device__ __align__(256) uint64_t scratch[kStaticScratchWords];

namespace cgapi {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), looked up on first use.
uint64_t* StaticScratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw ApiError{CGBDSM_ERR_INVALID, "device index beyond 64"};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        CGD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw ApiError{CGBDSM_ERR_DEVICE, "static scratch has no device address"};
        g_scratch_ptr[device] = (uint64_t*)p;
    }
    return g_scratch_ptr[device];
}

thread_local std::string g_err;

cgbdsm_status Fail(cgbdsm_status c, const std::string& m) {
    g_err = m;
    return c;
}

[[noreturn]] void Throw(cgbdsm_status c, const std::string& m) { throw ApiError{c, m}; }

// A new batch may be staged once the previous one is committed
// (cgbdsm_batch_update_global_index) or was skipped for an empty update
// index (main_calig.cpp:301-302 skips the rest of such a batch): staging over
// a batch still between apply and commit would drop that batch's commit.
void CheckCanStage(const cgbdsm_query* q) {
    if (q->phase == 0 || (q->phase == 3 && !q->u_nonempty)) return;
    static const char* names[] = {"idle", "staged", "applied", "update index built", "masks built", "matched"};
    Throw(CGBDSM_ERR_INVALID, std::string("the previous batch is ") + names[q->phase] +
                                  " but not committed: call cgbdsm_batch_update_global_index first");
}

void* DevAlloc(cgbdsm_ctx* ctx, size_t bytes) {
    if (!bytes) return nullptr;
    // Reserve the bytes against the budget before allocating, so that
    // queries allocating from several threads cannot pass the check together.
    const uint64_t before = ctx->used.fetch_add(bytes);
    if (ctx->budget && before + bytes > ctx->budget) {
        ctx->used.fetch_sub(bytes);
        Throw(CGBDSM_ERR_OOM, "device memory budget exceeded (" + std::to_string(before + bytes) + " > " +
                                  std::to_string(ctx->budget) + " bytes)");
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ctx->used.fetch_sub(bytes);
        Throw(CGBDSM_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return p;
}

void DevFree(cgbdsm_ctx* ctx, void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    uint64_t cur = ctx->used.load();
    while (!ctx->used.compare_exchange_weak(cur, cur - std::min<uint64_t>(cur, bytes))) {
    }
}

Tables MakeTables(cgbdsm_query* q) {
    Tables T;
    T.P = q->d_plan;
    T.m_rec = q->m_rec.p;
    T.u_rec = q->u_rec.p;
    T.pool = q->pool.p;
    T.u_vals = q->u_vals.p;
    T.masks = q->masks.p;
    T.ndir = q->ndir;
    T.qv = q->qv;
    T.qe = q->qe;
    T.dv = q->dv;
    T.W = q->W;
    return T;
}

void ReadCounters(cgbdsm_query* q) {
    CGD_HIP(hipMemcpyAsync(q->h_cnt, q->d_cnt, sizeof(Counters), hipMemcpyDeviceToHost, q->s));
    CGD_HIP(hipStreamSynchronize(q->s));
}

unsigned long long* Cnt(cgbdsm_query* q, unsigned long long Counters::*f) { return &(q->d_cnt->*f); }

void SetCounter(cgbdsm_query* q, unsigned long long Counters::*f, uint64_t v) {
    FillU64((uint64_t*)Cnt(q, f), v, 1, q->s);
}

// Several counters of q set by one launch on stream s.
void SetCounters(cgbdsm_query* q, std::initializer_list<std::pair<unsigned long long Counters::*, uint64_t>> fv,
                 hipStream_t s) {
    static const Counters probe{};
    WordSet ws{};
    for (const auto& [f, v] : fv) {
        if (ws.n == 16) Throw(CGBDSM_ERR_INVALID, "too many counters in one set");
        ws.off[ws.n] = (uint32_t)(((const char*)&(probe.*f) - (const char*)&probe) / sizeof(uint64_t));
        ws.val[ws.n++] = v;
    }
    SetWords((uint64_t*)q->d_cnt, ws, s);
}

void GrowDV(cgbdsm_query* q, uint32_t dv) {
    if (dv > q->dv_cap) {
        uint32_t nc = std::max<uint32_t>({dv, q->dv_cap + q->dv_cap / 2, 1024});
        const size_t old_rows = (size_t)q->dv_cap * q->ndir, new_rows = (size_t)nc * q->ndir;
        Ensure(q, q->m_rec, new_rows, true, old_rows);
        Ensure(q, q->u_rec, new_rows, true, old_rows);
        // zero the new tail (records are pairs / quads of u32)
        FillU64((uint64_t*)(q->m_rec.p + old_rows), 0, (q->m_rec.cap - old_rows) * 2, q->s);
        FillU64((uint64_t*)(q->u_rec.p + old_rows), 0, (q->u_rec.cap - old_rows), q->s);
        q->dv_cap = (uint32_t)std::min<size_t>(q->m_rec.cap, q->u_rec.cap) / q->ndir;
    }
    if (dv > q->dv) q->dv = dv;
    q->W = (q->dv + 63) / 64;
    q->vbits = BitsFor(std::max<uint32_t>(q->dv, 2));
    q->rbits = BitsFor((uint64_t)std::max<uint32_t>(q->dv, 1) * q->ndir);
}

void EnsureKeyScratch(cgbdsm_query* q, uint64_t n) {
    n = std::max<uint64_t>(n, 1);
    Ensure(q, q->k_alt, n);
    Ensure(q, q->flags, n + 1);
    Ensure(q, q->seg_row, n + 1);
    Ensure(q, q->seg_start, n + 2);
    Ensure(q, q->caps, n + 1);
    Ensure(q, q->sc_val, n);
    Ensure(q, q->sc_pos, n);
    Ensure(q, q->rs_hist, RadixHistWords(n));
    Ensure(q, q->scan_scratch, ScanScratchBytes(n + 2) + 64);
}

void CheckKeys(cgbdsm_query* q, const uint64_t* k, uint64_t n, uint32_t dv, const char* what) {
    if (n && !k) Throw(CGBDSM_ERR_INVALID, std::string(what) + ": null key pointer");
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t e = k[i] >> q->ushift, u = (k[i] >> 32) & q->umask, v = (uint32_t)k[i];
        if (e >= q->ndir || u >= dv || v >= dv)
            Throw(CGBDSM_ERR_INVALID, std::string(what) + " key " + std::to_string(i) + " out of range (edge " +
                                          std::to_string(e) + ", owner " + std::to_string(u) + ", target " +
                                          std::to_string(v) + ", dv " + std::to_string(dv) + ")");
    }
}

// Sort boundary keys `src` (n) into internal order in `dst` (dst may be any
// buffer of n u64; k_alt is used as the ping-pong partner unless it is dst).
int KBegin(cgbdsm_query* q, int cls);
void KEnd(cgbdsm_query* q, int idx);
void KAdd(cgbdsm_query* q, int cls, uint64_t bytes, uint64_t items);

void DecodeSort(cgbdsm_query* q, const uint64_t* src, uint64_t* dst, uint64_t* alt, uint64_t n) {
    const int k = KBegin(q, CGBDSM_K_SORT);
    LaunchDecode(src, dst, n, q->ushift, q->umask, q->ndir, q->vbits, q->s);
    RadixSortU64(dst, alt, n, q->vbits + q->rbits, q->rs_hist.p, q->s);
    KEnd(q, k);
    KAdd(q, CGBDSM_K_SORT, 0, n);
}

void GrowPool(cgbdsm_query* q, uint64_t min_cap) {
    ReadCounters(q);
    uint64_t used = q->h_cnt->pool_top;
    uint64_t nc = std::max<uint64_t>({min_cap, q->pool.cap * 2, used + (64ull << 20)});
    uint32_t* np = (uint32_t*)DevAlloc(q->ctx, nc * sizeof(uint32_t));
    if (q->pool.p && used)
        CGD_HIP(hipMemcpyAsync(np, q->pool.p, std::min<uint64_t>(used, q->pool.cap) * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, q->s));
    CGD_HIP(hipStreamSynchronize(q->s));
    DevFree(q->ctx, q->pool.p, q->pool.cap * sizeof(uint32_t));
    q->pool.p = np;
    q->pool.cap = nc;
}

// M <- M u keys (sorted internal keys, unique), retrying after pool growth.
void InsertSorted(cgbdsm_query* q, const uint64_t* keys, const unsigned long long* n_segs, uint64_t n_upper) {
    for (int attempt = 0;; attempt++) {
        SetCounter(q, &Counters::flags, 0);
        SetCounter(q, &Counters::bytes, 0);
        SetCounter(q, &Counters::n_long, 0);
        const int k = KBegin(q, CGBDSM_K_MERGE_INSERT);
        // q->flags (segment discovery's scan) is free again: long-list queue.
        LaunchMergeInsert(MakeTables(q), keys, q->vbits, q->seg_row.p, q->seg_start.p, n_segs, n_upper, q->sc_val.p,
                          q->sc_pos.p, q->flags.p, q->pool.cap, q->d_cnt, q->s);
        KEnd(q, k);
        ReadCounters(q);
        q->batch_idx_bytes += q->h_cnt->bytes;
        KAdd(q, CGBDSM_K_MERGE_INSERT, q->h_cnt->bytes, n_upper);
        if (!(q->h_cnt->flags & kFlagPool)) return;
        if (attempt >= 4) Throw(CGBDSM_ERR_OOM, "list pool overflow after growth");
        GrowPool(q, q->pool.cap * 2);
        // Lists that already merged see an empty B' and are skipped on retry.
    }
}

void ValidatePlanTables(const cgbdsm_plan* p, bool order) {
    auto bad = [](const std::string& m) { Throw(CGBDSM_ERR_INVALID, "plan: " + m); };
    if (p->qv < 3 || p->qv > (uint32_t)kMaxQV) bad("qv must be 3..12");
    if (p->qe < 2 || p->qe > (uint32_t)kMaxQE) bad("qe must be 2..42");
    if (p->ndir != 2 * p->qe) bad("ndir != 2*qe");
    if (p->ebits != 5 && p->ebits != 7) bad("ebits must be 5 or 7");
    if (p->ndir > (1u << p->ebits)) Throw(CGBDSM_ERR_LIMIT, "query has more directed edges than the key format holds");
    uint32_t seen = 0;
    for (uint32_t a = 0; a < p->qv; a++)
        for (uint32_t b = 0; b < p->qv; b++) {
            uint8_t e = p->eidx[a * kMaxQV + b];
            if (e == CGBDSM_NO_EDGE) continue;
            if (e >= p->ndir || a == b) bad("eidx out of range");
            if (p->dir_to_edge[e] >= p->qe) bad("dir_to_edge out of range");
            if (p->eidx[b * kMaxQV + a] == CGBDSM_NO_EDGE) bad("eidx not symmetric");
            seen++;
        }
    if (seen != p->ndir) bad("eidx does not list every directed edge");
    for (uint32_t i = 0; i < p->qe; i++) {
        if (p->qe_u[i] >= p->qv || p->qe_v[i] >= p->qv) bad("qe list out of range");
        if (p->eidx[p->qe_u[i] * kMaxQV + p->qe_v[i]] == CGBDSM_NO_EDGE) bad("qe list edge missing");
        if (p->dir_to_edge[p->eidx[p->qe_u[i] * kMaxQV + p->qe_v[i]]] != i) bad("dir_to_edge mismatch");
    }
    auto check_order = [&](const uint8_t* vs, const uint8_t* offs, const uint8_t* bni, int first_depth) {
        uint32_t mask = 0;
        for (uint32_t d = 0; d < p->qv; d++) {
            if (vs[d] >= p->qv || (mask >> vs[d]) & 1u) bad("order is not a permutation");
            mask |= 1u << vs[d];
        }
        for (uint32_t d = 0; d < p->qv; d++) {
            if (offs[d] > offs[d + 1] || offs[d + 1] > (uint32_t)kMaxDir) bad("bni offsets");
            if ((int)d < first_depth) continue;
            if (offs[d] == offs[d + 1]) bad("vertex without backward neighbour");
            for (uint32_t k = offs[d]; k < offs[d + 1]; k++) {
                if (bni[k] >= d) bad("backward neighbour not earlier");
                if (p->eidx[vs[bni[k]] * kMaxQV + vs[d]] == CGBDSM_NO_EDGE) bad("backward neighbour not adjacent");
            }
        }
    };
    for (uint32_t i = 0; i < p->qe; i++) {
        check_order(p->io_vs[i], p->io_bni_offs[i], p->io_bni[i], 2);
        if (p->io_vs[i][0] != p->qe_u[i] || p->io_vs[i][1] != p->qe_v[i]) bad("indexing order must start at the edge");
    }
    if (order) {
        check_order(p->go_vs, p->go_bni_offs, p->go_bni, 1);
        if (p->eidx[p->go_vs[0] * kMaxQV + p->go_vs[1]] == CGBDSM_NO_EDGE) bad("global order must start with an edge");
    }
}

void UploadPlan(cgbdsm_query* q) {
    CGD_HIP(hipMemcpyAsync(q->d_plan, &q->hplan, sizeof(DevPlan), hipMemcpyHostToDevice, q->s));
    CGD_HIP(hipStreamSynchronize(q->s));
}

EiList ActiveEdges(cgbdsm_query* q) {
    EiList L{};
    for (uint32_t e = 0; e < q->qe; e++)
        if (q->edge_ok[e]) L.ei[L.n++] = (uint8_t)e;
    return L;
}

hipEvent_t PoolEvent(cgbdsm_query* q) {
    if (!q->ev_pool.empty()) {
        hipEvent_t e = q->ev_pool.back();
        q->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    CGD_HIP(hipEventCreate(&e));
    return e;
}

int KBegin(cgbdsm_query* q, int cls) {
    if (!q->profile) return -1;
    cgbdsm_query::KPending k{cls, PoolEvent(q), PoolEvent(q)};
    CGD_HIP(hipEventRecord(k.a, q->s));
    q->kpend.push_back(k);
    return (int)q->kpend.size() - 1;
}

void KEnd(cgbdsm_query* q, int idx) {
    if (idx < 0) return;
    CGD_HIP(hipEventRecord(q->kpend[idx].b, q->s));
}

void KAdd(cgbdsm_query* q, int cls, uint64_t bytes, uint64_t items) {
    if (!q->profile) return;
    q->kt[cls].bytes += bytes;
    q->kt[cls].items += items;
}

void KHarvest(cgbdsm_query* q) {
    if (q->kpend.empty()) return;
    CGD_HIP(hipStreamSynchronize(q->s));
    for (auto& k : q->kpend) {
        float ms = 0;
        CGD_HIP(hipEventElapsedTime(&ms, k.a, k.b));
        q->kt[k.cls].ms += ms;
        q->kt[k.cls].launches++;
        q->ev_pool.push_back(k.a);
        q->ev_pool.push_back(k.b);
    }
    q->kpend.clear();
}

StealQueue MakeSteal(cgbdsm_query* q, bool donate) {
    StealQueue Q;
    Q.items = q->steal_items;
    Q.seq = q->steal_seq;
    Q.cap = q->steal_cap;
    Q.donate_min = donate ? q->donate_min : 0;
    Q.range_min = q->dfs_range_min;
    Q.abort = q->d_abort;
    Q.poll_ticks = std::max<long long>(1, q->ctx->wall_khz / 20);  // 50 us
    Q.host_ticks = (long long)q->ctx->wall_khz * 2;              // 2 ms
    return Q;
}

// Final DFS over the rows of the jobs in q->h_jobs (job j owns rows
// [h_rowbase[j], h_rowbase[j+1])), on q's stream with q's work queue; waits
// for it while honouring q's deadline: past the deadline the host raises the
// abort flag the kernel polls, and the call fails with CGBDSM_ERR_OVERFLOW
// (its counts would be incomplete).
void RunDfsJobs(cgbdsm_query* q, uint32_t* tuples, uint64_t tup_cap) {
    const uint64_t n = q->h_rowbase.back();
    const uint32_t depth = q->h_jobs.empty() ? 0 : q->h_jobs[0].D;
    LaunchQueueSync(q->d_cnt, q->s);
    {
        // flags, row cursor, pending work and idle count in one launch
        WordSet ws{};
        const size_t w[4] = {offsetof(Counters, flags), offsetof(Counters, q_rows), offsetof(Counters, q_pending),
                             offsetof(Counters, q_idle)};
        for (int i = 0; i < 4; i++) ws.off[i] = (uint32_t)(w[i] / sizeof(uint64_t));
        ws.val[2] = n;
        ws.n = 4;
        SetWords((uint64_t*)q->d_cnt, ws, q->s);
    }
    *(volatile uint32_t*)q->h_abort = 0;
    std::atomic_thread_fence(std::memory_order_seq_cst);
    if (!n) return;
    Ensure(q, q->d_jobs, q->h_jobs.size());
    Ensure(q, q->d_rowbase, q->h_rowbase.size());
    CGD_HIP(hipMemcpyAsync(q->d_jobs.p, q->h_jobs.data(), q->h_jobs.size() * sizeof(DfsJob), hipMemcpyHostToDevice,
                           q->s));
    CGD_HIP(hipMemcpyAsync(q->d_rowbase.p, q->h_rowbase.data(), q->h_rowbase.size() * sizeof(unsigned long long),
                           hipMemcpyHostToDevice, q->s));
    const int k = tuples ? -1 : KBegin(q, CGBDSM_K_DFS);
    uint32_t levels = 1;  // survivor stacks: qv - 1 levels cover every job
    for (const DfsJob& j : q->h_jobs) levels = std::max<uint32_t>(levels, j.T.qv - 1);
    {
        const uint32_t full = DfsLdsShape(levels, tuples != nullptr).stab_cap;
        q->st.last_dfs_lds_cache = q->dfs_cache_cap && q->dfs_cache_cap < full ? q->dfs_cache_cap & ~63u : full;
    }
    LaunchDfs(q->d_jobs.p, (uint32_t)q->h_jobs.size(), q->d_rowbase.p, n, q->d_cnt, MakeSteal(q, tuples == nullptr),
              tuples, tup_cap, q->ctx->num_cus, q->dfs_bpc, q->dfs_grab, levels, q->dfs_cache_cap, q->s);
    KEnd(q, k);
    CGD_HIP(hipGetLastError());
    if (q->deadline_ms > 0) {
        const auto t0 = std::chrono::steady_clock::now();
        for (;;) {
            hipError_t e = hipStreamQuery(q->s);
            if (e == hipSuccess) break;
            if (e != hipErrorNotReady) CGD_HIP(e);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (ms > q->deadline_ms) {
                *(volatile uint32_t*)q->h_abort = 1;
                std::atomic_thread_fence(std::memory_order_seq_cst);
                CGD_HIP(hipStreamSynchronize(q->s));
                *(volatile uint32_t*)q->h_abort = 0;
                ReadCounters(q);
                unsigned long long part = 0;
                for (uint32_t e = 0; e < q->qe; e++) part += q->h_cnt->per_edge[e];
                Throw(CGBDSM_ERR_OVERFLOW, "final DFS exceeded the deadline of " + std::to_string(q->deadline_ms) +
                                               " ms; count incomplete (partial " + std::to_string(part) +
                                               " matches, " + std::to_string(q->h_cnt->scanned) + " candidates, " +
                                               std::to_string(q->h_cnt->probes) + " probes, " + std::to_string(n) +
                                               " rows at depth " + std::to_string(depth) + ", " +
                                               std::to_string(q->h_cnt->q_rows.v) + " rows taken, " +
                                               std::to_string(q->h_cnt->donated) + " donated; pending " +
                                               std::to_string(q->h_cnt->q_pending.v) + ", idle " +
                                               std::to_string(q->h_cnt->q_idle.v) + ", queue " +
                                               std::to_string(q->h_cnt->q_head.v) + "/" +
                                               std::to_string(q->h_cnt->q_tail.v) + ", flags " +
                                               std::to_string(q->h_cnt->flags) + ")");
            }
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    }
}

DfsJob MakeJob(cgbdsm_query* q, const uint32_t* rows, uint32_t depth) {
    return DfsJob{MakeTables(q), rows, q->d_cnt, depth,
                  (q->dfs_shortest ? kJobShortest : 0u) | (q->dfs_suffix ? kJobSuffix : 0u) |
                      (q->dfs_wsearch ? kJobWaveSearch | (q->dfs_wsearch << kJobWaveShift) : 0u)};
}

void RunDfs(cgbdsm_query* q, const uint32_t* rows, uint64_t n, uint32_t depth, uint32_t* tuples, uint64_t tup_cap) {
    q->h_jobs.assign(1, MakeJob(q, rows, depth));
    q->h_rowbase.assign({0ull, (unsigned long long)n});
    RunDfsJobs(q, tuples, tup_cap);
}

float Elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}


// ---------------------------------------------------------------- GPMA layout
void PmaAlloc(cgbdsm_query* q, uint32_t nleaves) {
    auto& P = q->pma;
    const uint64_t slots = (uint64_t)nleaves * kPmaLeaf;
    Ensure(q, P.slot, slots);
    Ensure(q, P.shadow, slots);
    Ensure(q, P.lo, nleaves);
    Ensure(q, P.S, 2ull * nleaves);
    Ensure(q, P.cnt, nleaves + 1ull);
    Ensure(q, P.ins_cnt, nleaves);
    Ensure(q, P.win_of, nleaves);
    Ensure(q, P.gpos, nleaves + 1ull);
    Ensure(q, P.wins, nleaves);
    Ensure(q, P.meta, nleaves);
    Ensure(q, P.chosen, (2ull * nleaves + 3) / 4);
    P.nleaves = nleaves;
    P.height = 0;
    while ((1u << P.height) < nleaves) P.height++;
}

uint32_t PmaLeavesFor(uint64_t n) {
    // Bulk load / rebuild at about half density.
    const uint64_t want = std::max<uint64_t>(1, (2 * n + kPmaLeaf - 1) / kPmaLeaf);
    uint32_t l = 1;
    while (l < want) l <<= 1;
    return l;
}

// Even spread of n sorted PMA keys in P.keys over a fresh array.
void PmaSpread(cgbdsm_query* q, uint64_t n) {
    auto& P = q->pma;
    PmaAlloc(q, PmaLeavesFor(n));
    LaunchPmaSpread(P.keys.p, n, P.slot.p, P.cnt.p, P.lo.p, P.nleaves, q->s);
    FillU32(P.cnt.p + P.nleaves, 0, 1, q->s);
    P.n = n;
}

// Bulk load from n sorted unique index keys (device encoding).
void PmaBulk(cgbdsm_query* q, const uint64_t* keys, uint64_t n) {
    auto& P = q->pma;
    Ensure(q, P.keys, std::max<uint64_t>(n, 1));
    LaunchPmaEncode(keys, n, q->vbits, P.keys.p, q->s);
    PmaSpread(q, n);
}

// Flat lists (M for matching) from the array: one streaming pass.
// may_grow = false: a pool too small for the current entries fails with
// CGBDSM_ERR_LIMIT instead of being reallocated (nothing is changed then).
void PmaExtract(cgbdsm_query* q, bool may_grow = true) {
    auto& P = q->pma;
    const int k = KBegin(q, CGBDSM_K_GPMA_EXTRACT);
    ScanU32(P.cnt.p, P.gpos.p, P.nleaves + 1ull, (uint64_t*)Cnt(q, &Counters::n_segs), q->scan_scratch.p, q->s);
    ReadCounters(q);
    const uint64_t n = q->h_cnt->n_segs;
    if (q->pool.cap < std::max<uint64_t>(n, 1)) {
        if (!may_grow) {
            KEnd(q, k);
            Throw(CGBDSM_ERR_LIMIT, "GPMA flat lists are stale and the list pool must grow to refresh them");
        }
        Release(q, q->pool);
        const uint64_t want = n + n / 2 + (16ull << 20);
        q->pool.p = (uint32_t*)DevAlloc(q->ctx, want * sizeof(uint32_t));
        q->pool.cap = want;
    }
    LaunchPmaFlat(P.slot.p, P.cnt.p, P.gpos.p, P.nleaves, P.shadow.p, q->s);
    const uint64_t rows = (uint64_t)q->dv * q->ndir;
    LaunchPmaRows(P.shadow.p, n, q->pool.p, q->m_rec.p, rows, q->s);
    KEnd(q, k);
    KAdd(q, CGBDSM_K_GPMA_EXTRACT, 8ull * (uint64_t)P.nleaves * kPmaLeaf + 12ull * n + 16ull * rows, n);
    P.n = n;
    P.stale = false;
    q->st.main_entries = n;
}

// Rebuild at a larger size with `extra` sorted absent PMA keys merged in
// (the whole array overflowed its root density bound).
void PmaRebuild(cgbdsm_query* q, const uint64_t* extra, uint64_t n_extra) {
    auto& P = q->pma;
    ScanU32(P.cnt.p, P.gpos.p, P.nleaves + 1ull, (uint64_t*)Cnt(q, &Counters::n_segs), q->scan_scratch.p, q->s);
    ReadCounters(q);
    const uint64_t n_old = q->h_cnt->n_segs, n = n_old + n_extra;
    // The extra keys live in P.keys2: copy them out before P.keys grows.
    DBuf<uint64_t> tmp;
    tmp.p = (uint64_t*)DevAlloc(q->ctx, std::max<uint64_t>(n_extra, 1) * sizeof(uint64_t));
    tmp.cap = std::max<uint64_t>(n_extra, 1);
    if (n_extra) CGD_HIP(hipMemcpyAsync(tmp.p, extra, n_extra * 8, hipMemcpyDeviceToDevice, q->s));
    Ensure(q, P.keys, std::max<uint64_t>(n, 1));
    LaunchPmaFlat(P.slot.p, P.cnt.p, P.gpos.p, P.nleaves, P.keys.p, q->s);
    if (n_extra) CGD_HIP(hipMemcpyAsync(P.keys.p + n_old, tmp.p, n_extra * 8, hipMemcpyDeviceToDevice, q->s));
    CGD_HIP(hipStreamSynchronize(q->s));
    DevFree(q->ctx, tmp.p, tmp.cap * sizeof(uint64_t));
    Ensure(q, P.keys2, std::max<uint64_t>(n, 1));
    Ensure(q, q->rs_hist, RadixHistWords(n));
    const uint32_t bits = kPmaVBits + BitsFor((uint64_t)q->dv_cap * q->ndir + 1);
    RadixSortU64(P.keys.p, P.keys2.p, n, bits, q->rs_hist.p, q->s);
    PmaSpread(q, n);
}

// M <- M u keys: nk sorted unique index keys (device encoding).
void PmaInsert(cgbdsm_query* q, const uint64_t* keys, uint64_t nk) {
    auto& P = q->pma;
    if (!nk) return;
    const int k = KBegin(q, CGBDSM_K_GPMA_UPDATE);
    Ensure(q, P.keys, nk);
    Ensure(q, P.keys2, nk);
    Ensure(q, P.leaf, nk);
    Ensure(q, P.leaf2, nk);
    Ensure(q, P.keep, nk + 1);
    Ensure(q, P.pos, nk + 1);
    Ensure(q, q->scan_scratch, ScanScratchBytes(std::max<uint64_t>(nk + 2, P.nleaves + 2ull)) + 64);
    LaunchPmaEncode(keys, nk, q->vbits, P.keys.p, q->s);
    LaunchPmaRoute(P.keys.p, nk, P.lo.p, P.nleaves, P.leaf.p, q->s);
    LaunchPmaAbsent(P.keys.p, nk, P.leaf.p, P.slot.p, P.cnt.p, P.keep.p, q->s);
    unsigned long long* n_new = Cnt(q, &Counters::n_segs);
    ScanU32(P.keep.p, P.pos.p, nk, (uint64_t*)n_new, q->scan_scratch.p, q->s);
    LaunchPmaCompact(P.keys.p, P.leaf.p, nk, P.keep.p, P.pos.p, P.keys2.p, P.leaf2.p, q->s);
    SetCounter(q, &Counters::flags, 0);
    LaunchPmaWindows(P.cnt.p, P.ins_cnt.p, P.leaf2.p, n_new, nk, P.S.p, (uint8_t*)P.chosen.p, P.nleaves, P.height,
                     q->d_cnt, q->s);
    ReadCounters(q);
    const uint64_t nn = q->h_cnt->n_segs;
    if (nn && (q->h_cnt->flags & kFlagPool)) {
        KEnd(q, k);
        PmaRebuild(q, P.keys2.p, nn);
        q->st.gpma_rebuilds++;
    } else if (nn) {
        ScanU32(P.cnt.p, P.gpos.p, P.nleaves + 1ull, (uint64_t*)Cnt(q, &Counters::pool_top), q->scan_scratch.p,
                q->s);
        SetCounter(q, &Counters::rows_out, 0);
        LaunchPmaRebalance((const uint8_t*)P.chosen.p, P.nleaves, P.height, P.wins.p, Cnt(q, &Counters::rows_out),
                           P.win_of.p, P.gpos.p, P.leaf2.p, P.keys2.p, n_new, nk, P.meta.p, P.slot.p, P.shadow.p,
                           P.cnt.p, P.lo.p, Cnt(q, &Counters::pma_level), q->s);
        KEnd(q, k);
        ReadCounters(q);
        q->st.gpma_windows += q->h_cnt->rows_out;
        q->st.gpma_top_level = std::max<uint64_t>(q->st.gpma_top_level, q->h_cnt->pma_level);
    } else {
        KEnd(q, k);
    }
    KAdd(q, CGBDSM_K_GPMA_UPDATE, 8ull * nk, nk);
    P.n += nn;
}

// M <- M \\ keys: n sorted index keys (device encoding).
void PmaDelete(cgbdsm_query* q, const uint64_t* keys, uint64_t n) {
    auto& P = q->pma;
    if (!n) return;
    const int k = KBegin(q, CGBDSM_K_GPMA_UPDATE);
    Ensure(q, P.keys, n);
    Ensure(q, P.leaf, n);
    LaunchPmaEncode(keys, n, q->vbits, P.keys.p, q->s);
    LaunchPmaRoute(P.keys.p, n, P.lo.p, P.nleaves, P.leaf.p, q->s);
    LaunchPmaDelete(P.keys.p, P.leaf.p, n, P.slot.p, P.cnt.p, q->s);
    KEnd(q, k);
    KAdd(q, CGBDSM_K_GPMA_UPDATE, 8ull * n, n);
}

void IndexInitDevice(cgbdsm_query* q, uint64_t n, uint32_t dv_count, uint32_t* card, float* avg) {
    SplitInvalidate(q);
    GrowDV(q, std::max<uint32_t>(dv_count, 1));
    q->dv = dv_count;
    q->W = (q->dv + 63) / 64;
    FillU64((uint64_t*)q->m_rec.p, 0, q->m_rec.cap * 2, q->s);
    FillU64((uint64_t*)q->u_rec.p, 0, q->u_rec.cap, q->s);
    FillU64((uint64_t*)q->d_cnt, 0, sizeof(Counters) / 8, q->s);
    q->u_rows_upper = 0;
    EnsureKeyScratch(q, n);
    DecodeSort(q, q->k_ins.p, q->k_alt.p, q->k_ins.p, n);
    LaunchUnique(q->k_alt.p, q->k_ins.p, n, q->flags.p, q->scan_scratch.p, Cnt(q, &Counters::n_unique), q->s);
    LaunchSegments(q->k_ins.p, Cnt(q, &Counters::n_unique), n, q->vbits, q->flags.p, q->seg_row.p, q->seg_start.p,
                   Cnt(q, &Counters::n_segs), q->scan_scratch.p, q->s);
    LaunchInitCaps(q->seg_start.p, Cnt(q, &Counters::n_segs), n, q->caps.p, q->s);
    // An empty initial index (no lit pair) still scans one capacity: zero it.
    if (!n) FillU64(q->caps.p, 0, 1, q->s);
    // exclusive scan of capacities -> offsets; total into pool_top
    // (flags hold the head positions needed by the layout kernel)
    ScanU64(q->caps.p, q->caps.p, std::max<uint64_t>(n, 1), (uint64_t*)Cnt(q, &Counters::pool_top),
            q->scan_scratch.p, q->s);
    ReadCounters(q);
    const uint64_t total = q->h_cnt->pool_top;
    const uint64_t want = std::max<uint64_t>(total + total / 2 + (16ull << 20), 1ull << 24);
    if (q->pool.cap < want) {
        Release(q, q->pool);
        q->pool.p = (uint32_t*)DevAlloc(q->ctx, want * sizeof(uint32_t));
        q->pool.cap = want;
    }
    LaunchInitLayout(MakeTables(q), q->k_ins.p, Cnt(q, &Counters::n_unique), n, q->vbits, q->flags.p,
                     q->seg_row.p, q->seg_start.p, Cnt(q, &Counters::n_segs), q->caps.p, q->d_cnt, q->s);
    ReadCounters(q);
    for (uint32_t e = 0; e < q->ndir; e++) {
        const uint64_t c = q->h_cnt->card[e], nd = q->h_cnt->nodes[e];
        if (card) card[e] = (uint32_t)c;
        if (avg) avg[e] = nd ? (float)(uint32_t)c / (float)(uint32_t)nd : 0.0f;
    }
    q->st.main_entries = q->h_cnt->n_unique;
    if (q->layout == CGBDSM_LAYOUT_GPMA) {
        // The lists built above give the cardinalities; the matching
        // lists come from the array from here on.
        PmaBulk(q, q->k_ins.p, q->h_cnt->n_unique);
        PmaExtract(q);
    }
    q->inited = true;
    q->phase = 0;
}

void StageDeviceKeys(cgbdsm_query* q, uint64_t n_ins, uint64_t n_del, uint64_t n_upd, const uint8_t* edge_ok,
                     uint32_t dv) {
    if (dv > (1u << (32 - q->ebits)) || dv > q->hplan.v_mask)
        Throw(CGBDSM_ERR_LIMIT, "compressed vertex count exceeds the key / row-tag width");
    GrowDV(q, dv);
    const uint64_t nmax = std::max<uint64_t>({n_ins, n_del, n_upd, 1});
    Ensure(q, q->k_ins, std::max<uint64_t>(n_ins, 1), true, n_ins);
    Ensure(q, q->k_del, std::max<uint64_t>(n_del, 1), true, n_del);
    Ensure(q, q->k_upd, std::max<uint64_t>(n_upd, 1), true, n_upd);
    Ensure(q, q->k_usort, std::max<uint64_t>(n_upd, 1));
    Ensure(q, q->u_vals, std::max<uint64_t>(n_upd, 1));
    Ensure(q, q->u_rows, std::max<uint64_t>(n_upd, 1), true, q->u_rows_upper);
    EnsureKeyScratch(q, nmax);
    q->n_ins = n_ins;
    q->n_del = n_del;
    q->n_upd = n_upd;
    memcpy(q->edge_ok, edge_ok, q->qe);
    q->st.last_ins = n_ins;
    q->st.last_del = n_del;
    q->st.last_upd = n_upd;
    q->phase = 1;
}

}  // namespace cgapi

extern "C" {

const char* cgbdsm_last_error(void) { return g_err.c_str(); }

const char* cgbdsm_version(void) { return "cgbdsm 0.1 gfx950 wave64 keys:ebits5|7"; }

cgbdsm_status cgbdsm_ctx_create(int device, uint64_t budget_bytes, cgbdsm_ctx** out) {
    return Guard([&] {
        if (!out) Throw(CGBDSM_ERR_INVALID, "null out");
        int n = 0;
        CGD_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) Throw(CGBDSM_ERR_INVALID, "no such device");
        CGD_HIP(hipSetDevice(device));
        hipDeviceProp_t pr;
        CGD_HIP(hipGetDeviceProperties(&pr, device));
        auto* c = new cgbdsm_ctx;
        c->device = device;
        c->budget = budget_bytes;
        c->num_cus = pr.multiProcessorCount > 0 ? pr.multiProcessorCount : 256;
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            c->wall_khz = khz;
        try {
            StaticScratch(device);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

cgbdsm_status cgbdsm_ctx_destroy(cgbdsm_ctx* ctx) {
    return Guard([&] { delete ctx; });
}

cgbdsm_status cgbdsm_ctx_memory(cgbdsm_ctx* ctx, uint64_t* used_bytes, uint64_t* budget_bytes) {
    return Guard([&] {
        if (!ctx) Throw(CGBDSM_ERR_INVALID, "null context");
        if (used_bytes) *used_bytes = ctx->used.load();
        if (budget_bytes) *budget_bytes = ctx->budget;
    });
}

cgbdsm_status cgbdsm_query_create(cgbdsm_ctx* ctx, const cgbdsm_plan* plan, cgbdsm_query** out) {
    return Guard([&] {
        if (!ctx || !plan || !out) Throw(CGBDSM_ERR_INVALID, "null argument");
        ValidatePlanTables(plan, false);
        CGD_HIP(hipSetDevice(ctx->device));
        auto* q = new cgbdsm_query;
        try {
            q->ctx = ctx;
            q->plan = *plan;
            q->qv = plan->qv;
            q->qe = plan->qe;
            q->ndir = plan->ndir;
            q->ebits = plan->ebits;
            q->ushift = 64 - plan->ebits;
            q->umask = (1u << (32 - plan->ebits)) - 1u;
            CGD_HIP(hipStreamCreateWithFlags(&q->s, hipStreamNonBlocking));
            for (auto& e : q->ev) CGD_HIP(hipEventCreate(&e));
            CGD_HIP(hipEventCreateWithFlags(&q->fr_ev, hipEventDisableTiming));
            q->d_plan = (DevPlan*)DevAlloc(ctx, sizeof(DevPlan));
            q->d_cnt = (Counters*)DevAlloc(ctx, sizeof(Counters));
            CGD_HIP(hipHostMalloc((void**)&q->h_cnt, sizeof(Counters), hipHostMallocDefault));
            memset(q->h_cnt, 0, sizeof(Counters));
            CGD_HIP(hipHostMalloc((void**)&q->h_abort, 64, hipHostMallocMapped | hipHostMallocCoherent));
            *(volatile uint32_t*)q->h_abort = 0;
            CGD_HIP(hipHostGetDevicePointer((void**)&q->d_abort, q->h_abort, 0));
            q->steal_items = (StealItem*)DevAlloc(ctx, q->steal_cap * sizeof(StealItem));
            q->steal_seq = (unsigned long long*)DevAlloc(ctx, q->steal_cap * sizeof(unsigned long long));
            FillU64((uint64_t*)q->steal_seq, 0, q->steal_cap, q->s);
            FillU64((uint64_t*)q->d_cnt, 0, sizeof(Counters) / 8, q->s);
            q->hplan.p = *plan;
            q->hplan.tag_shift = q->qe <= 32 ? 27 : 26;
            q->hplan.v_mask = (1u << q->hplan.tag_shift) - 1u;
            for (uint32_t e = 0; e < q->qe; e++) {
                q->hplan.sa[e] = plan->eidx[plan->qe_u[e] * kMaxQV + plan->qe_v[e]];
                q->hplan.sb[e] = plan->eidx[plan->qe_v[e] * kMaxQV + plan->qe_u[e]];
                for (uint32_t v = 0; v < (uint32_t)kMaxQV; v++) q->hplan.all_ones[e][v] = 0;
            }
            UploadPlan(q);
        } catch (...) {
            cgbdsm_query_destroy(q);
            throw;
        }
        *out = q;
    });
}

cgbdsm_status cgbdsm_query_destroy(cgbdsm_query* q) {
    return Guard([&] {
        if (!q) return;
        if (q->s) (void)hipStreamSynchronize(q->s);
        SplitFree(q);
        S1Detach(q);
        Release(q, q->m_rec);
        Release(q, q->u_rec);
        Release(q, q->pool);
        for (auto* b : {&q->k_ins, &q->k_del, &q->k_upd, &q->k_usort, &q->k_alt, &q->seg_start, &q->caps, &q->masks})
            Release(q, *b);
        for (auto* b : {&q->flags, &q->seg_row, &q->sc_val, &q->sc_pos, &q->rs_hist, &q->u_vals, &q->u_rows,
                        &q->front[0], &q->front[1]})
            Release(q, *b);
        Release(q, q->scan_scratch);
        for (auto* b : {&q->pma.slot, &q->pma.shadow, &q->pma.lo, &q->pma.S, &q->pma.keys, &q->pma.keys2})
            Release(q, *b);
        for (auto* b : {&q->pma.cnt, &q->pma.ins_cnt, &q->pma.win_of, &q->pma.gpos, &q->pma.wins, &q->pma.leaf,
                        &q->pma.leaf2, &q->pma.keep, &q->pma.pos, &q->pma.chosen})
            Release(q, *b);
        Release(q, q->pma.meta);
        Release(q, q->d_jobs);
        Release(q, q->d_rowbase);
        Release(q, q->tag_hist);
        for (auto& r : q->fifo) {
            Release(q, r.ins);
            Release(q, r.del);
            Release(q, r.upd);
        }
        if (q->d_plan) DevFree(q->ctx, q->d_plan, sizeof(DevPlan));
        if (q->d_cnt) DevFree(q->ctx, q->d_cnt, sizeof(Counters));
        if (q->h_cnt) (void)hipHostFree(q->h_cnt);
        if (q->h_abort) (void)hipHostFree(q->h_abort);
        if (q->steal_items) DevFree(q->ctx, q->steal_items, q->steal_cap * sizeof(StealItem));
        if (q->steal_seq) DevFree(q->ctx, q->steal_seq, q->steal_cap * sizeof(unsigned long long));
        for (auto& e : q->ev)
            if (e) (void)hipEventDestroy(e);
        if (q->fr_ev) (void)hipEventDestroy(q->fr_ev);
        for (auto& k : q->kpend) {
            (void)hipEventDestroy(k.a);
            (void)hipEventDestroy(k.b);
        }
        for (auto& e : q->ev_pool) (void)hipEventDestroy(e);
        if (q->s) (void)hipStreamDestroy(q->s);
        delete q;
    });
}

cgbdsm_status cgbdsm_query_set_plan(cgbdsm_query* q, const cgbdsm_plan* plan) {
    return Guard([&] {
        if (!q || !plan) Throw(CGBDSM_ERR_INVALID, "null argument");
        if (plan->qv != q->qv || plan->qe != q->qe || plan->ndir != q->ndir || plan->ebits != q->ebits ||
            memcmp(plan->eidx, q->plan.eidx, sizeof(plan->eidx)) != 0)
            Throw(CGBDSM_ERR_INVALID, "plan does not belong to this query");
        ValidatePlanTables(plan, true);
        q->plan = *plan;
        q->hplan.p = *plan;
        for (uint32_t e = 0; e < q->qe; e++)
            for (uint32_t v = 0; v < q->qv; v++) q->hplan.all_ones[e][v] = plan->rebuild_v[e][v] ? 0 : 1;
        // Product suffix: the longest tail of the order whose positions are
        // mutually non-adjacent (every backward neighbour precedes the tail),
        // as Plan::GenerateCartesianProductInfoGlobal (graph/plan.cpp:899-952)
        // classifies it, shortened until no label class in it has more than
        // three members (the closed forms cover classes of up to 3).
        const uint32_t qv = plan->qv, last = qv - 1;
        uint32_t cps = qv;
        for (uint32_t st = last; st >= 2; st--) {
            bool ok = true;
            for (uint32_t j = st; j <= last && ok; j++)
                for (uint32_t k = plan->go_bni_offs[j]; k < plan->go_bni_offs[j + 1]; k++)
                    if (plan->go_bni[k] >= st) ok = false;
            if (!ok) break;
            cps = st;
        }
        for (; cps < qv; cps++) {
            uint32_t worst = 0;
            for (uint32_t j = cps; j <= last; j++) {
                uint32_t m = 0;
                for (uint32_t i = cps; i <= last; i++) m += plan->go_lcls[i] == plan->go_lcls[j];
                worst = std::max(worst, m);
            }
            if (worst <= 3) break;
        }
        q->hplan.cp_start = cps;
        q->hplan.sfx_batch = 0;
        if (cps < qv) {
            bool singles = true;
            for (uint32_t j = cps; j < qv; j++)
                for (uint32_t i = j + 1; i < qv; i++)
                    if (plan->go_lcls[i] == plan->go_lcls[j]) singles = false;
            q->hplan.sfx_batch = singles ? 1u : 0u;
        }
        // The suffix position with the most backward neighbours (two or more)
        // keeps the intersection of its lists from all but the deepest one in
        // LDS; it only changes when one of those shallower positions does.
        q->hplan.stab_pos = 0xFF;
        q->hplan.stab_var_k = 0;
        q->hplan.stab_deps = 0;
        {
            uint32_t best = 1;
            for (uint32_t j = cps; j < qv; j++) {
                const uint32_t nbn = plan->go_bni_offs[j + 1] - plan->go_bni_offs[j];
                if (nbn > best) {
                    best = nbn;
                    q->hplan.stab_pos = j;
                }
            }
            if (q->hplan.stab_pos != 0xFF) {
                const uint32_t j = q->hplan.stab_pos;
                uint32_t mb = 0, mk = 0;
                for (uint32_t k = plan->go_bni_offs[j]; k < plan->go_bni_offs[j + 1]; k++)
                    if (plan->go_bni[k] >= mb) {
                        mb = plan->go_bni[k];
                        mk = k - plan->go_bni_offs[j];
                    }
                q->hplan.stab_var_k = mk;
                for (uint32_t k = plan->go_bni_offs[j]; k < plan->go_bni_offs[j + 1]; k++)
                    if (plan->go_bni[k] != mb) q->hplan.stab_deps |= 1u << plan->go_bni[k];
            }
        }
        for (uint32_t j = 0; j < (uint32_t)kMaxQV; j++) {
            uint32_t mb = 0;
            if (j < qv)
                for (uint32_t k = plan->go_bni_offs[j]; k < plan->go_bni_offs[j + 1]; k++)
                    mb = std::max<uint32_t>(mb, plan->go_bni[k]);
            q->hplan.maxbn[j] = (uint8_t)mb;
        }
        for (uint32_t x = 0; x < (uint32_t)kMaxQV; x++) {
            uint32_t m = 0;
            for (uint32_t j = 0; j < qv && x < qv; j++)
                for (uint32_t k = plan->go_bni_offs[j]; k < plan->go_bni_offs[j + 1]; k++)
                    if (plan->go_bni[k] == x) m |= 1u << j;
            q->hplan.dep_mask[x] = m;
        }
        UploadPlan(q);
        q->have_order = true;
    });
}

cgbdsm_status cgbdsm_query_set_layout(cgbdsm_query* q, int layout) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        if (layout != CGBDSM_LAYOUT_GAMMA && layout != CGBDSM_LAYOUT_GPMA) Throw(CGBDSM_ERR_INVALID, "unknown layout");
        if (q->inited) Throw(CGBDSM_ERR_INVALID, "the layout is chosen before cgbdsm_index_init");
        q->layout = layout;
    });
}

cgbdsm_status cgbdsm_index_init(cgbdsm_query* q, const uint64_t* keys, uint64_t n, uint32_t dv_count,
                                uint32_t* cardinalities, float* avg_degrees) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        if (q->s1) Throw(CGBDSM_ERR_INVALID, "the device Stage 1 built this query's index (cgbdsm_calig_attach)");
        if (dv_count > (1u << (32 - q->ebits)) || dv_count > q->hplan.v_mask)
            Throw(CGBDSM_ERR_LIMIT, "compressed vertex count exceeds the key / row-tag width");
        CheckKeys(q, keys, n, dv_count, "index_init");
        Ensure(q, q->k_ins, std::max<uint64_t>(n, 1));
        if (n) CGD_HIP(hipMemcpyAsync(q->k_ins.p, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice, q->s));
        IndexInitDevice(q, n, dv_count, cardinalities, avg_degrees);
    });
}

cgbdsm_status cgbdsm_batch_stage(cgbdsm_query* q, const uint64_t* ins, uint64_t n_ins, const uint64_t* del,
                                 uint64_t n_del, const uint64_t* upd, uint64_t n_upd, const uint8_t* edge_ok,
                                 uint32_t dv_count) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        if (!q->have_order) Throw(CGBDSM_ERR_INVALID, "cgbdsm_query_set_plan must run before the first batch");
        if (q->s1) Throw(CGBDSM_ERR_INVALID, "this query's batches come from cgbdsm_calig_update");
        CheckCanStage(q);
        if (!edge_ok) Throw(CGBDSM_ERR_INVALID, "null edge_ok");
        if (dv_count < q->dv) Throw(CGBDSM_ERR_INVALID, "compressed vertex count cannot shrink");
        if (dv_count > (1u << (32 - q->ebits)) || dv_count > q->hplan.v_mask)
            Throw(CGBDSM_ERR_LIMIT, "compressed vertex count exceeds the key / row-tag width");
        CheckKeys(q, ins, n_ins, dv_count, "ins");
        CheckKeys(q, del, n_del, dv_count, "del");
        CheckKeys(q, upd, n_upd, dv_count, "upd");
        GrowDV(q, dv_count);
        const uint64_t nmax = std::max<uint64_t>({n_ins, n_del, n_upd, 1});
        Ensure(q, q->k_ins, std::max<uint64_t>(n_ins, 1));
        Ensure(q, q->k_del, std::max<uint64_t>(n_del, 1));
        Ensure(q, q->k_upd, std::max<uint64_t>(n_upd, 1));
        Ensure(q, q->k_usort, std::max<uint64_t>(n_upd, 1));
        Ensure(q, q->u_vals, std::max<uint64_t>(n_upd, 1));
        // keep the previous batch's touched-row list readable until U is rebuilt
        Ensure(q, q->u_rows, std::max<uint64_t>(n_upd, 1), true, q->u_rows_upper);
        EnsureKeyScratch(q, nmax);
        if (n_ins) CGD_HIP(hipMemcpyAsync(q->k_ins.p, ins, n_ins * 8, hipMemcpyHostToDevice, q->s));
        if (n_del) CGD_HIP(hipMemcpyAsync(q->k_del.p, del, n_del * 8, hipMemcpyHostToDevice, q->s));
        if (n_upd) CGD_HIP(hipMemcpyAsync(q->k_upd.p, upd, n_upd * 8, hipMemcpyHostToDevice, q->s));
        q->n_ins = n_ins;
        q->n_del = n_del;
        q->n_upd = n_upd;
        memcpy(q->edge_ok, edge_ok, q->qe);
        q->st.last_ins = n_ins;
        q->st.last_del = n_del;
        q->st.last_upd = n_upd;
        q->phase = 1;
    });
}

cgbdsm_status cgbdsm_batch_enqueue(cgbdsm_query* q, const uint64_t* ins, uint64_t n_ins, const uint64_t* del,
                                   uint64_t n_del, const uint64_t* upd, uint64_t n_upd, const uint8_t* edge_ok,
                                   uint32_t dv_count) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        if (q->s1) Throw(CGBDSM_ERR_INVALID, "this query's batches come from cgbdsm_calig_update");
        if (!edge_ok) Throw(CGBDSM_ERR_INVALID, "null edge_ok");
        const uint32_t prev = q->fifo.empty() ? q->dv : q->fifo.back().dv;
        if (dv_count < prev) Throw(CGBDSM_ERR_INVALID, "compressed vertex count cannot shrink");
        if (dv_count > (1u << (32 - q->ebits)) || dv_count > q->hplan.v_mask)
            Throw(CGBDSM_ERR_LIMIT, "compressed vertex count exceeds the key / row-tag width");
        CheckKeys(q, ins, n_ins, dv_count, "ins");
        CheckKeys(q, del, n_del, dv_count, "del");
        CheckKeys(q, upd, n_upd, dv_count, "upd");
        cgbdsm_query::Resident r;
        try {
            Ensure(q, r.ins, std::max<uint64_t>(n_ins, 1));
            Ensure(q, r.del, std::max<uint64_t>(n_del, 1));
            Ensure(q, r.upd, std::max<uint64_t>(n_upd, 1));
            if (n_ins) CGD_HIP(hipMemcpyAsync(r.ins.p, ins, n_ins * 8, hipMemcpyHostToDevice, q->s));
            if (n_del) CGD_HIP(hipMemcpyAsync(r.del.p, del, n_del * 8, hipMemcpyHostToDevice, q->s));
            if (n_upd) CGD_HIP(hipMemcpyAsync(r.upd.p, upd, n_upd * 8, hipMemcpyHostToDevice, q->s));
            CGD_HIP(hipStreamSynchronize(q->s));
        } catch (...) {
            Release(q, r.ins);
            Release(q, r.del);
            Release(q, r.upd);
            throw;
        }
        r.n_ins = n_ins;
        r.n_del = n_del;
        r.n_upd = n_upd;
        memcpy(r.edge_ok, edge_ok, q->qe);
        r.dv = dv_count;
        q->fifo.push_back(r);
    });
}

cgbdsm_status cgbdsm_batch_stage_next(cgbdsm_query* q, int* staged) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        if (!q->have_order) Throw(CGBDSM_ERR_INVALID, "cgbdsm_query_set_plan must run before the first batch");
        if (staged) *staged = 0;
        if (q->fifo.empty()) return;
        CheckCanStage(q);
        cgbdsm_query::Resident& r = q->fifo.front();
        GrowDV(q, r.dv);
        const uint64_t nmax = std::max<uint64_t>({r.n_ins, r.n_del, r.n_upd, 1});
        Ensure(q, q->k_usort, std::max<uint64_t>(r.n_upd, 1));
        Ensure(q, q->u_vals, std::max<uint64_t>(r.n_upd, 1));
        Ensure(q, q->u_rows, std::max<uint64_t>(r.n_upd, 1), true, q->u_rows_upper);
        EnsureKeyScratch(q, nmax);
        std::swap(q->k_ins, r.ins);
        std::swap(q->k_del, r.del);
        std::swap(q->k_upd, r.upd);
        q->n_ins = r.n_ins;
        q->n_del = r.n_del;
        q->n_upd = r.n_upd;
        memcpy(q->edge_ok, r.edge_ok, q->qe);
        q->st.last_ins = r.n_ins;
        q->st.last_del = r.n_del;
        q->st.last_upd = r.n_upd;
        Release(q, r.ins);
        Release(q, r.del);
        Release(q, r.upd);
        q->fifo.pop_front();
        q->phase = 1;
        if (staged) *staged = 1;
    });
}

cgbdsm_status cgbdsm_transfer_update_all(cgbdsm_query* q) {
    return Guard([&] {
        if (!q || q->phase != 1) Throw(CGBDSM_ERR_INVALID, "transfer_update_all needs a staged batch");
        // The host mirror takes the batch's keys before the device's sort
        // reuses their buffers.
        if (q->split) SplitOnApply(q);
        SetCounter(q, &Counters::bytes, 0);
        SetCounter(q, &Counters::rows_touched, 0);
        q->batch_idx_bytes = 0;
        if (q->layout == CGBDSM_LAYOUT_GPMA) {
            if (q->n_ins) {
                DecodeSort(q, q->k_ins.p, q->k_alt.p, q->k_ins.p, q->n_ins);
                LaunchUnique(q->k_alt.p, q->k_ins.p, q->n_ins, q->flags.p, q->scan_scratch.p,
                             Cnt(q, &Counters::n_unique), q->s);
                ReadCounters(q);
                PmaInsert(q, q->k_ins.p, q->h_cnt->n_unique);
            }
            if (q->n_del) {
                DecodeSort(q, q->k_del.p, q->k_alt.p, q->k_del.p, q->n_del);
                PmaDelete(q, q->k_alt.p, q->n_del);
            }
            PmaExtract(q);
            q->phase = 2;
            return;
        }
        if (q->n_ins) {
            DecodeSort(q, q->k_ins.p, q->k_alt.p, q->k_ins.p, q->n_ins);
            LaunchUnique(q->k_alt.p, q->k_ins.p, q->n_ins, q->flags.p, q->scan_scratch.p, Cnt(q, &Counters::n_unique),
                         q->s);
            LaunchSegments(q->k_ins.p, Cnt(q, &Counters::n_unique), q->n_ins, q->vbits, q->flags.p, q->seg_row.p,
                           q->seg_start.p, Cnt(q, &Counters::n_segs), q->scan_scratch.p, q->s);
            InsertSorted(q, q->k_ins.p, Cnt(q, &Counters::n_segs), q->n_ins);
        }
        if (q->n_del) {
            DecodeSort(q, q->k_del.p, q->k_alt.p, q->k_del.p, q->n_del);
            SetCounter(q, &Counters::n_unique, q->n_del);
            LaunchSegments(q->k_alt.p, Cnt(q, &Counters::n_unique), q->n_del, q->vbits, q->flags.p, q->seg_row.p,
                           q->seg_start.p, Cnt(q, &Counters::n_segs), q->scan_scratch.p, q->s);
            SetCounter(q, &Counters::bytes, 0);
            SetCounter(q, &Counters::n_long, 0);
            const int k = KBegin(q, CGBDSM_K_MERGE_REMOVE);
            LaunchMergeRemove(MakeTables(q), q->k_alt.p, q->vbits, q->seg_row.p, q->seg_start.p,
                              Cnt(q, &Counters::n_segs), q->n_del, q->flags.p, q->d_cnt, q->s);
            KEnd(q, k);
            ReadCounters(q);
            q->batch_idx_bytes += q->h_cnt->bytes;
            KAdd(q, CGBDSM_K_MERGE_REMOVE, q->h_cnt->bytes, q->n_del);
        }
        q->st.last_idx_bytes = q->batch_idx_bytes;
        q->phase = 2;
    });
}

cgbdsm_status cgbdsm_transfer_update_index(cgbdsm_query* q, int* nonempty) {
    return Guard([&] {
        if (!q || q->phase != 2) Throw(CGBDSM_ERR_INVALID, "transfer_update_index must follow transfer_update_all");
        if (q->split) SplitOnUpdateIndex(q);
        LaunchClearU(q->u_rec.p, q->u_rows.p, Cnt(q, &Counters::u_rows), q->u_rows_upper, q->s);
        SetCounter(q, &Counters::u_rows, 0);
        q->u_rows_upper = 0;
        q->u_nonempty = q->n_upd > 0;
        if (q->n_upd) {
            DecodeSort(q, q->k_upd.p, q->k_usort.p, q->k_alt.p, q->n_upd);
            SetCounter(q, &Counters::u_dups, 0);
            LaunchCountDups(q->k_usort.p, q->n_upd, q->d_cnt, q->s);
            SetCounter(q, &Counters::n_unique, q->n_upd);
            const int k = KBegin(q, CGBDSM_K_UPDATE_INDEX);
            LaunchSegments(q->k_usort.p, Cnt(q, &Counters::n_unique), q->n_upd, q->vbits, q->flags.p, q->seg_row.p,
                           q->seg_start.p, Cnt(q, &Counters::n_segs), q->scan_scratch.p, q->s);
            LaunchBuildU(q->u_rec.p, q->k_usort.p, q->n_upd, q->vbits, q->u_vals.p, q->seg_row.p, q->seg_start.p,
                         Cnt(q, &Counters::n_segs), q->u_rows.p, Cnt(q, &Counters::u_rows), q->s);
            KEnd(q, k);
            KAdd(q, CGBDSM_K_UPDATE_INDEX, 12ull * q->n_upd, q->n_upd);
            q->u_rows_upper = q->n_upd;
        }
        if (nonempty) *nonempty = q->u_nonempty ? 1 : 0;
        q->phase = 3;
    });
}

cgbdsm_status cgbdsm_build_local_index(cgbdsm_query* q) {
    return Guard([&] {
        if (!q || q->phase != 3) Throw(CGBDSM_ERR_INVALID, "build_local_index must follow transfer_update_index");
        const EiList L = ActiveEdges(q);
        Ensure(q, q->masks, std::max<uint64_t>((uint64_t)q->qe * kMaxQV * q->W, 1));
        const Tables T = MakeTables(q);
        const int kt = KBegin(q, CGBDSM_K_MASKS);
        LaunchMaskClear(q->masks.p, (uint64_t)kMaxQV * q->W, L, q->s);
        EiSeedInfo info{};
        for (uint32_t k = 0; k < L.n; k++) {
            const uint32_t ei = L.ei[k];
            const uint8_t a = q->plan.io_vs[ei][0], b = q->plan.io_vs[ei][1];
            info.ei[info.n] = (uint8_t)ei;
            info.i01[info.n] = q->plan.eidx[a * kMaxQV + b];
            info.i10[info.n] = q->plan.eidx[b * kMaxQV + a];
            info.vs0[info.n] = a;
            info.vs1[info.n] = b;
            info.n++;
        }
        LaunchMaskSeed(T, q->u_rows.p, Cnt(q, &Counters::u_rows), q->u_rows_upper, info, q->s);
        if (q->profile) SetCounter(q, &Counters::bytes, 0);
        uint64_t tiles = 0;
        for (uint32_t d = 2; d < q->qv; d++) {
            EiList D{};
            for (uint32_t k = 0; k < L.n; k++) {
                const uint32_t ei = L.ei[k];
                if (q->plan.rebuild_v[ei][q->plan.io_vs[ei][d]]) D.ei[D.n++] = (uint8_t)ei;
            }
            LaunchMaskDepth(T, d, D, q->d_cnt, q->s);
            tiles += D.n;
        }
        KEnd(q, kt);
        if (q->profile) {
            // Tile writes: ceil(dv/8) bytes per rebuilt tile and per seeded tile.
            ReadCounters(q);
            KAdd(q, CGBDSM_K_MASKS, q->h_cnt->bytes + (tiles + 2ull * L.n) * ((q->dv + 7) / 8), tiles);
        }
        q->phase = 4;
        q->fr_ready = false;
    });
}

cgbdsm_status cgbdsm_set_match_params(cgbdsm_query* q, uint64_t bfs_row_limit, uint64_t frontier_cap,
                                      uint32_t donate_min) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        q->bfs_row_limit = bfs_row_limit;
        q->frontier_cap_rows = frontier_cap;
        q->donate_min = donate_min;
    });
}

cgbdsm_status cgbdsm_set_dfs_tuning(cgbdsm_query* q, uint32_t rows_per_grab, uint32_t blocks_per_cu) {
    return Guard([&] {
        if (!q || rows_per_grab > 1024) Throw(CGBDSM_ERR_INVALID, "bad DFS tuning");
        q->dfs_grab = rows_per_grab ? rows_per_grab : 4;
        q->dfs_bpc = blocks_per_cu;
    });
}

cgbdsm_status cgbdsm_set_dfs_search(cgbdsm_query* q, int wave_search) {
    return Guard([&] {
        // 64 buckets of at least one entry each: lists of 64 entries or more.
        if (!q || wave_search < 0 || (wave_search > 1 && (wave_search < 64 || (wave_search & (wave_search - 1)))))
            Throw(CGBDSM_ERR_INVALID, "wave_search must be 0, 1 or a power of two >= 64");
        uint32_t lg = 0;
        if (wave_search == 1) lg = 8;  // lists of 256 entries or more
        while (wave_search > 1 && (1 << lg) < wave_search) lg++;
        q->dfs_wsearch = (uint32_t)lg;
    });
}

cgbdsm_status cgbdsm_set_dfs_steal(cgbdsm_query* q, uint32_t range_min, uint32_t cache_cap) {
    return Guard([&] {
        if (!q || range_min > (1u << 20)) Throw(CGBDSM_ERR_INVALID, "bad DFS stealing knobs");
        q->dfs_range_min = range_min ? std::max<uint32_t>(range_min, 64) : 1024;
        q->dfs_cache_cap = cache_cap;
    });
}

cgbdsm_status cgbdsm_set_dfs_options(cgbdsm_query* q, int shortest_list, int product_suffix) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        q->dfs_shortest = shortest_list != 0;
        q->dfs_suffix = product_suffix != 0;
    });
}

cgbdsm_status cgbdsm_set_deadline_ms(cgbdsm_query* q, double ms) {
    return Guard([&] {
        if (!q || ms < 0) Throw(CGBDSM_ERR_INVALID, "bad deadline");
        q->deadline_ms = ms;
    });
}

// Seeds and BFS levels of one query's matching (phase 4): the rows left for
// the final DFS, their depth and the buffer holding them, kept in q->fr.
// The last kernel (the tag grouping) may still run on q's stream on return;
// q->fr_ev marks the frontier's completion there, for a batched DFS launch
// on another query's stream to wait on.
static const Frontier& BuildFrontier(cgbdsm_query* q) {
    if (q->fr_ready) return q->fr;
    const EiList L = ActiveEdges(q);
    SetCounters(q, {{&Counters::rows_out, 0}, {&Counters::flags, 0}, {&Counters::tuples, 0}}, q->s);
    const Tables T = MakeTables(q);
    const bool fixed_cap = q->frontier_cap_rows != 0;
    uint64_t cap_rows = fixed_cap ? q->frontier_cap_rows : (1u << 20);  // seeds grow the buffer on overflow
    Ensure(q, q->front[0], cap_rows * 2);
    // Product suffix from position 2 (counted for all siblings at once):
    // depth-1 rows, so that the DFS scans position 1 itself and the sizes
    // that depend on v0 alone are computed once per v0, not once per (v0, v1).
    bool seed1 = false;
    if (q->hplan.sfx_batch && q->dfs_suffix && q->hplan.cp_start == 2 && q->qv > 2) {
        ReadCounters(q);
        seed1 = q->h_cnt->u_dups == 0;
    }
    const uint32_t words = seed1 ? 1u : 2u;
    // Step 1: seeds (retry once with a buffer of the exact size on overflow).
    uint64_t n = 0;
    for (int attempt = 0;; attempt++) {
        const int ks = KBegin(q, CGBDSM_K_SEED);
        if (seed1) LaunchSeed1(T, L, q->front[0].p, q->front[0].cap, q->d_cnt, q->s);
        else LaunchSeed(T, L, q->front[0].p, q->front[0].cap / 2, q->d_cnt, q->s);
        KEnd(q, ks);
        ReadCounters(q);
        n = q->h_cnt->rows_out;
        if (!(q->h_cnt->flags & kFlagFrontier)) break;
        if (attempt) Throw(CGBDSM_ERR_OVERFLOW, "seed frontier overflow");
        Ensure(q, q->front[0], n * words + 64);
        SetCounter(q, &Counters::rows_out, 0);
        SetCounter(q, &Counters::flags, 0);
    }
    uint32_t depth = seed1 ? 1u : 2u, cur = 0;
    // Seeds come out grouped by owner; group them by edge tag.
    if (n > 1 && q->qe > 1) {
        Ensure(q, q->front[1], n * words);
        Ensure(q, q->tag_hist, 2 * kMaxQE);
        FillU64((uint64_t*)q->tag_hist.p, 0, 2 * kMaxQE, q->s);
        const int kg = KBegin(q, CGBDSM_K_SEED);
        LaunchGroupByTag(q->front[0].p, n, words, q->hplan.tag_shift, q->tag_hist.p, q->front[1].p, q->s);
        KEnd(q, kg);
        cur = 1;
    }
    // Step 2: BFS levels while the frontier is small.
    // With a batched product suffix the DFS must still scan the position
    // before the suffix (cp_start - 1) itself: BFS levels stop short of it.
    const bool batched_suffix =
        q->hplan.sfx_batch && q->dfs_suffix && q->h_cnt->u_dups == 0 && q->hplan.cp_start < q->qv;
    while (n > 0 && n < q->bfs_row_limit && depth + 1 < q->qv &&
           !(batched_suffix && depth + 2 > q->hplan.cp_start)) {
        const uint64_t want_rows =
            fixed_cap ? q->frontier_cap_rows : std::min<uint64_t>(std::max<uint64_t>(n * 64, 1u << 20), 1ull << 28);
        const uint64_t want_words = want_rows * (depth + 1);
        Ensure(q, q->front[cur ^ 1], want_words);
        SetCounter(q, &Counters::rows_out, 0);
        SetCounter(q, &Counters::flags, 0);
        const int ke = KBegin(q, CGBDSM_K_EXTEND);
        LaunchExtend(T, q->front[cur].p, n, depth, q->front[cur ^ 1].p, q->front[cur ^ 1].cap / (depth + 1),
                     q->d_cnt, q->s);
        KEnd(q, ke);
        ReadCounters(q);
        KAdd(q, CGBDSM_K_EXTEND, 0, n);
        if (q->h_cnt->flags & kFlagFrontier) break;  // keep this level, count it by DFS
        n = q->h_cnt->rows_out;
        cur ^= 1;
        depth++;
    }
    CGD_HIP(hipEventRecord(q->fr_ev, q->s));
    q->fr = Frontier{n, depth, cur};
    q->fr_ready = true;
    return q->fr;
}

// Zeroes the per-edge counts (and the frontier/tuple flags) of `q` on the
// stream `s` that will run the final DFS.
static void ResetMatchCounters(cgbdsm_query* q, hipStream_t s) {
    static_assert(offsetof(Counters, per_edge) == offsetof(Counters, count_ovf) + 8, "count_ovf precedes per_edge");
    FillU64((uint64_t*)&q->d_cnt->count_ovf, 0, 1 + kMaxQE, s);
    FillU64((uint64_t*)&q->d_cnt->tuples, 0, 1, s);
}

static void ResetDfsCounters(cgbdsm_query* q) {
    SetCounters(q, {{&Counters::scanned, 0}, {&Counters::probes, 0}, {&Counters::donated, 0},
                    {&Counters::wave_ticks, 0}, {&Counters::idle_ticks, 0}, {&Counters::steps, 0},
                    {&Counters::step_lanes, 0}, {&Counters::batched, 0}}, q->s);
}

// Statistics of the final-DFS launch whose queue lived in `leader`.
static void CheckDfs(cgbdsm_query* leader, uint64_t rows) {
    ReadCounters(leader);
    const Counters* c = leader->h_cnt;
    if (c->flags & kFlagBadRow)
        Throw(CGBDSM_ERR_DEVICE, "final DFS read a frontier row with an out-of-range tag or vertex id");
    if (c->flags & (kFlagSteal | kFlagAborted))
        Throw(CGBDSM_ERR_OVERFLOW, "final DFS did not complete (flags " + std::to_string(c->flags) + ")");
    if (rows) KAdd(leader, CGBDSM_K_DFS, 4ull * (c->scanned + c->probes), rows);
}

// Per-query results of a completed final DFS; `launch` holds the counters of
// the launch's work queue (the query itself for a single-query launch).
// A query whose batch count did not fit 64 bits (the count is then wrong:
// the call fails instead of returning it).
static void CheckWide(cgbdsm_query* const* qs, int nq) {
    std::string names;
    for (int i = 0; i < nq; i++)
        if (qs[i]->count_wide) names += (names.empty() ? "" : ", ") + std::to_string(i);
    if (!names.empty())
        Throw(CGBDSM_ERR_OVERFLOW, "the batch's match count does not fit 64 bits (query " + names + " of the call)");
}

static void FinishMatching(cgbdsm_query* q, const Frontier& f, const cgbdsm_query* launch,
                           unsigned long long* num_matches, uint64_t* per_edge) {
    const Counters* c = launch->h_cnt;
    q->st.last_dfs_scanned = c->scanned;
    q->st.last_dfs_probes = c->probes;
    q->st.last_dfs_donated = c->donated;
    q->st.last_dfs_steps = c->steps;
    q->st.last_dfs_step_lanes = c->step_lanes;
    q->st.last_dfs_batched = c->batched;
    q->st.last_dfs_busy = c->wave_ticks ? 1.0 - (double)c->idle_ticks / (double)c->wave_ticks : 0.0;
    unsigned long long total = 0;
    bool wide = q->h_cnt->count_ovf != 0;
    for (uint32_t e = 0; e < q->qe; e++) {
        const unsigned long long pe = q->h_cnt->per_edge[e];
        wide |= total + pe < total;
        total += pe;
        if (per_edge) per_edge[e] = pe;
    }
    q->count_wide = wide;
    if (num_matches) *num_matches += total;
    q->last_depth = f.depth;
    q->last_rows = f.n;
    q->last_buf = f.cur;
    q->st.last_frontier_rows = f.n;
    q->st.last_frontier_depth = f.depth;
    q->st.last_matches = total;
    q->phase = 5;
}

cgbdsm_status cgbdsm_build_frontier(cgbdsm_query* q) {
    return Guard([&] {
        if (!q || q->phase != 4) Throw(CGBDSM_ERR_INVALID, "build_frontier must follow build_local_index");
        BuildFrontier(q);
    });
}

cgbdsm_status cgbdsm_matching(cgbdsm_query* q, unsigned long long* num_matches, uint64_t* per_edge) {
    return Guard([&] {
        if (!q || (q->phase != 4 && q->phase != 5)) Throw(CGBDSM_ERR_INVALID, "matching must follow build_local_index");
        const Frontier f = BuildFrontier(q);
        // Step 5: final DFS count from the current level.
        ResetMatchCounters(q, q->s);
        ResetDfsCounters(q);
        const uint64_t k = q->split ? SplitPlan(q, f.n) : 0;
        if (k) {
            // Rows [0, k) are counted on the host while the GPU counts the rest.
            SplitStart(q, q->front[f.cur].p, k, f.n, f.depth);
            try {
                CGD_HIP(hipEventRecord(q->ev[6], q->s));
                RunDfs(q, q->front[f.cur].p + k * f.depth, f.n - k, f.depth, nullptr, 0);
                CGD_HIP(hipEventRecord(q->ev[7], q->s));
                CheckDfs(q, f.n - k);
            } catch (...) {
                SplitJoin(q, true);
                throw;
            }
            SplitFinish(q, Elapsed(q->ev[6], q->ev[7]));
        } else {
            RunDfs(q, q->front[f.cur].p, f.n, f.depth, nullptr, 0);
            CheckDfs(q, f.n);
        }
        FinishMatching(q, f, q, num_matches, per_edge);
        CheckWide(&q, 1);
    });
}

cgbdsm_status cgbdsm_matching_batch(cgbdsm_query* const* qs, int nq, unsigned long long* num_matches,
                                    uint64_t* const* per_edge) {
    return Guard([&] {
        if (!qs || nq <= 0 || nq > CGBDSM_MAX_BATCH_QUERIES) Throw(CGBDSM_ERR_INVALID, "bad query batch");
        for (int i = 0; i < nq; i++) {
            if (!qs[i] || (qs[i]->phase != 4 && qs[i]->phase != 5))
                Throw(CGBDSM_ERR_INVALID, "matching must follow build_local_index");
            if (qs[i]->ctx != qs[0]->ctx) Throw(CGBDSM_ERR_INVALID, "batched queries must share one context");
            for (int k = 0; k < i; k++)
                if (qs[k] == qs[i]) Throw(CGBDSM_ERR_INVALID, "query listed twice in a batch");
        }
        // Seeds and BFS levels, one query after another, for queries whose
        // frontier cgbdsm_build_frontier has not built already.
        std::vector<Frontier> fr(nq);
        for (int i = 0; i < nq; i++) fr[i] = BuildFrontier(qs[i]);
        // One final-DFS launch for all of them, on the first query's stream
        // and work queue, after every other query's frontier is complete on
        // its own stream (the tag grouping may still be running there).
        cgbdsm_query* lead = qs[0];
        for (int i = 1; i < nq; i++) CGD_HIP(hipStreamWaitEvent(lead->s, qs[i]->fr_ev, 0));
        lead->h_jobs.clear();
        lead->h_rowbase.assign(1, 0ull);
        for (int i = 0; i < nq; i++) {
            if (!fr[i].n) continue;
            lead->h_jobs.push_back(MakeJob(qs[i], qs[i]->front[fr[i].cur].p, fr[i].depth));
            lead->h_rowbase.push_back(lead->h_rowbase.back() + fr[i].n);
        }
        // Per-query counts are zeroed on the launch's stream (the grouping
        // kernels still in flight on the other streams do not touch them).
        for (int i = 0; i < nq; i++) ResetMatchCounters(qs[i], lead->s);
        ResetDfsCounters(lead);
        RunDfsJobs(lead, nullptr, 0);
        // The other queries' counters come back on the launch's stream, with
        // the lead's (CheckDfs): one wait instead of one per query.
        for (int i = 0; i < nq; i++)
            if (qs[i] != lead)
                CGD_HIP(hipMemcpyAsync(qs[i]->h_cnt, qs[i]->d_cnt, sizeof(Counters), hipMemcpyDeviceToHost, lead->s));
        CheckDfs(lead, lead->h_rowbase.back());
        for (int i = 0; i < nq; i++) {
            FinishMatching(qs[i], fr[i], lead, num_matches ? &num_matches[i] : nullptr,
                           per_edge ? per_edge[i] : nullptr);
        }
        CheckWide(qs, nq);
    });
}

cgbdsm_status cgbdsm_match_tuples(cgbdsm_query* q, uint32_t* out, uint64_t cap, uint64_t* n_out, uint8_t* ei_out) {
    return Guard([&] {
        if (!q || q->phase != 5) Throw(CGBDSM_ERR_INVALID, "match_tuples must follow matching");
        const uint64_t words = (q->qv + 1) * std::max<uint64_t>(cap, 1);
        uint32_t* d_out = (uint32_t*)DevAlloc(q->ctx, words * 4);
        try {
            SetCounter(q, &Counters::tuples, 0);
            SetCounter(q, &Counters::flags, 0);
            FillU64((uint64_t*)&q->d_cnt->per_edge[0], 0, kMaxQE, q->s);
            RunDfs(q, q->front[q->last_buf].p, q->last_rows, q->last_depth, d_out, cap);
            ReadCounters(q);
            if (q->h_cnt->flags & kFlagTuples) Throw(CGBDSM_ERR_OVERFLOW, "tuple capacity exceeded");
            const uint64_t n = q->h_cnt->tuples;
            if (n_out) *n_out = n;
            const uint64_t k = std::min(n, cap);
            std::vector<uint32_t> h(k * (q->qv + 1));
            if (k) CGD_HIP(hipMemcpy(h.data(), d_out, h.size() * 4, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < k; i++) {
                if (out) memcpy(out + i * q->qv, &h[i * (q->qv + 1)], q->qv * 4);
                if (ei_out) ei_out[i] = (uint8_t)h[i * (q->qv + 1) + q->qv];
            }
        } catch (...) {
            DevFree(q->ctx, d_out, words * 4);
            throw;
        }
        DevFree(q->ctx, d_out, words * 4);
    });
}

cgbdsm_status cgbdsm_batch_update_global_index(cgbdsm_query* q) {
    return Guard([&] {
        if (!q || (q->phase != 5 && q->phase != 4 && q->phase != 3))
            Throw(CGBDSM_ERR_INVALID, "batch_update_global_index must follow transfer_update_index");
        if (q->split) SplitOnCommit(q);
        if (q->n_upd) {
            LaunchUnique(q->k_usort.p, q->k_alt.p, q->n_upd, q->flags.p, q->scan_scratch.p, Cnt(q, &Counters::n_unique),
                         q->s);
            if (q->layout == CGBDSM_LAYOUT_GPMA) {
                ReadCounters(q);
                PmaInsert(q, q->k_alt.p, q->h_cnt->n_unique);
                q->pma.stale = true;
            } else {
                LaunchSegments(q->k_alt.p, Cnt(q, &Counters::n_unique), q->n_upd, q->vbits, q->flags.p, q->seg_row.p,
                               q->seg_start.p, Cnt(q, &Counters::n_segs), q->scan_scratch.p, q->s);
                InsertSorted(q, q->k_alt.p, Cnt(q, &Counters::n_segs), q->n_upd);
            }
        }
        q->st.batches++;
        q->phase = 0;
    });
}

cgbdsm_status cgbdsm_run_batch(cgbdsm_query* q, unsigned long long* num_matches, uint64_t* per_edge, int* skipped) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        auto chk = [](cgbdsm_status s) {
            if (s != CGBDSM_OK) Throw(s, g_err);
        };
        CGD_HIP(hipEventRecord(q->ev[0], q->s));
        chk(cgbdsm_transfer_update_all(q));
        CGD_HIP(hipEventRecord(q->ev[1], q->s));
        int ne = 0;
        chk(cgbdsm_transfer_update_index(q, &ne));
        CGD_HIP(hipEventRecord(q->ev[2], q->s));
        if (skipped) *skipped = ne ? 0 : 1;
        if (!ne) {
            q->phase = 0;
            q->st.batches++;
            CGD_HIP(hipStreamSynchronize(q->s));
            q->st.ms_apply = Elapsed(q->ev[0], q->ev[1]);
            q->st.ms_update_index = Elapsed(q->ev[1], q->ev[2]);
            q->st.ms_masks = q->st.ms_match = q->st.ms_commit = 0;
            return;
        }
        chk(cgbdsm_build_local_index(q));
        CGD_HIP(hipEventRecord(q->ev[3], q->s));
        chk(cgbdsm_matching(q, num_matches, per_edge));
        CGD_HIP(hipEventRecord(q->ev[4], q->s));
        chk(cgbdsm_batch_update_global_index(q));
        CGD_HIP(hipEventRecord(q->ev[5], q->s));
        CGD_HIP(hipStreamSynchronize(q->s));
        q->st.ms_apply = Elapsed(q->ev[0], q->ev[1]);
        q->st.ms_update_index = Elapsed(q->ev[1], q->ev[2]);
        q->st.ms_masks = Elapsed(q->ev[2], q->ev[3]);
        q->st.ms_match = Elapsed(q->ev[3], q->ev[4]);
        q->st.ms_commit = Elapsed(q->ev[4], q->ev[5]);
        KHarvest(q);
    });
}

cgbdsm_status cgbdsm_profile(cgbdsm_query* q, int enable) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        KHarvest(q);
        q->profile = enable != 0;
    });
}

cgbdsm_status cgbdsm_kernel_times(cgbdsm_query* q, cgbdsm_ktime* out, int reset) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        CGD_HIP(hipStreamSynchronize(q->s));
        KHarvest(q);
        if (out) memcpy(out, q->kt, sizeof(q->kt));
        if (reset) memset(q->kt, 0, sizeof(q->kt));
    });
}

cgbdsm_status cgbdsm_index_export(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        if (which == 0 && q->layout == CGBDSM_LAYOUT_GPMA && q->pma.stale) PmaExtract(q);
        const uint64_t nrows = (uint64_t)q->dv * q->ndir;
        uint64_t* lens = (uint64_t*)DevAlloc(q->ctx, (nrows + 1) * 8);
        void* scr = DevAlloc(q->ctx, ScanScratchBytes(nrows + 1) + 64);
        uint64_t* tot = (uint64_t*)DevAlloc(q->ctx, 8);
        uint64_t total = 0;
        std::vector<uint64_t> h;
        try {
            LaunchRowLens(MakeTables(q), nrows, which, lens, q->s);
            ScanU64(lens, lens, nrows, tot, scr, q->s);
            CGD_HIP(hipMemcpyAsync(&total, tot, 8, hipMemcpyDeviceToHost, q->s));
            CGD_HIP(hipStreamSynchronize(q->s));
            if (total) {
                uint64_t* d = (uint64_t*)DevAlloc(q->ctx, total * 8);
                LaunchExport(MakeTables(q), nrows, which, lens, q->ushift, d, q->s);
                h.resize(total);
                CGD_HIP(hipMemcpyAsync(h.data(), d, total * 8, hipMemcpyDeviceToHost, q->s));
                CGD_HIP(hipStreamSynchronize(q->s));
                DevFree(q->ctx, d, total * 8);
            }
        } catch (...) {
            DevFree(q->ctx, lens, (nrows + 1) * 8);
            DevFree(q->ctx, scr, ScanScratchBytes(nrows + 1) + 64);
            DevFree(q->ctx, tot, 8);
            throw;
        }
        DevFree(q->ctx, lens, (nrows + 1) * 8);
        DevFree(q->ctx, scr, ScanScratchBytes(nrows + 1) + 64);
        DevFree(q->ctx, tot, 8);
        std::sort(h.begin(), h.end());
        if (n_out) *n_out = total;
        if (out) memcpy(out, h.data(), std::min(cap, total) * 8);
    });
}

cgbdsm_status cgbdsm_index_export_static(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out) {
    return Guard([&] {
        if (!q || (which != 0 && which != 1)) Throw(CGBDSM_ERR_INVALID, "bad query or index selector");
        uint64_t* base = StaticScratch(q->ctx->device);
        // Refreshing stale GPMA flat lists must not grow the list pool.
        if (which == 0 && q->layout == CGBDSM_LAYOUT_GPMA && q->pma.stale) PmaExtract(q, false);
        const uint64_t nrows = (uint64_t)q->dv * q->ndir;
        // Scratch layout (u64 words): row lengths then offsets [nrows + 1],
        // the scan total, the scan's tile sums, then the exported keys.
        const uint64_t w_total = nrows + 1;
        const uint64_t w_scan = w_total + 1;
        const uint64_t scan_words = (ScanScratchBytes(nrows + 1) + 64 + 7) / 8;
        const uint64_t w_keys = (w_scan + scan_words + 31) & ~31ull;
        if (w_keys > kStaticScratchWords)
            Throw(CGBDSM_ERR_LIMIT, "index has " + std::to_string(nrows) + " rows; the static export holds at most " +
                                        std::to_string(kStaticScratchWords * 8) + " bytes of scratch");
        uint64_t* lens = base;
        uint64_t* tot = base + w_total;
        LaunchRowLens(MakeTables(q), nrows, which, lens, q->s);
        ScanU64(lens, lens, nrows, tot, base + w_scan, q->s);
        uint64_t total = 0;
        CGD_HIP(hipMemcpyAsync(&total, tot, 8, hipMemcpyDeviceToHost, q->s));
        CGD_HIP(hipStreamSynchronize(q->s));
        if (!nrows) total = 0;
        if (n_out) *n_out = total;
        if (w_keys + total > kStaticScratchWords)
            Throw(CGBDSM_ERR_LIMIT, "index has " + std::to_string(total) + " entries; the static export holds at most " +
                                        std::to_string(kStaticScratchWords - w_keys) + " for " +
                                        std::to_string(nrows) + " rows");
        if (total && (!out || cap < total))
            Throw(CGBDSM_ERR_INVALID, "out holds " + std::to_string(cap) + " keys, the index has " +
                                          std::to_string(total) + " (*n_out is set)");
        if (total) {
            LaunchExport(MakeTables(q), nrows, which, lens, q->ushift, base + w_keys, q->s);
            CGD_HIP(hipMemcpyAsync(out, base + w_keys, total * 8, hipMemcpyDeviceToHost, q->s));
            CGD_HIP(hipStreamSynchronize(q->s));
            std::sort(out, out + total);
        }
    });
}

cgbdsm_status cgbdsm_staged_keys(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out) {
    return Guard([&] {
        if (!q || which < 0 || which > 2) Throw(CGBDSM_ERR_INVALID, "bad query or key selector");
        if (q->phase != 1) Throw(CGBDSM_ERR_INVALID, "no batch is staged");
        const DBuf<uint64_t>& b = which == 0 ? q->k_ins : which == 1 ? q->k_del : q->k_upd;
        const uint64_t n = which == 0 ? q->n_ins : which == 1 ? q->n_del : q->n_upd;
        if (n_out) *n_out = n;
        if (out && cap && n) {
            CGD_HIP(hipMemcpyAsync(out, b.p, std::min(cap, n) * 8, hipMemcpyDeviceToHost, q->s));
            CGD_HIP(hipStreamSynchronize(q->s));
        }
    });
}

cgbdsm_status cgbdsm_mask_export(cgbdsm_query* q, uint32_t ei, uint32_t qv, uint64_t* words, uint64_t nwords,
                                 int* all_ones) {
    return Guard([&] {
        if (!q || ei >= q->qe || qv >= q->qv) Throw(CGBDSM_ERR_INVALID, "bad tile");
        if (q->phase < 4) Throw(CGBDSM_ERR_INVALID, "masks are built by build_local_index");
        const bool ones = q->hplan.all_ones[ei][qv] != 0;
        if (all_ones) *all_ones = ones ? 1 : 0;
        if (!ones && words && nwords) {
            if (!q->edge_ok[ei]) Throw(CGBDSM_ERR_INVALID, "tile of an edge without batch lists is not built");
            CGD_HIP(hipMemcpyAsync(words, q->masks.p + ((uint64_t)ei * kMaxQV + qv) * q->W,
                                   std::min<uint64_t>(nwords, q->W) * 8, hipMemcpyDeviceToHost, q->s));
            CGD_HIP(hipStreamSynchronize(q->s));
        }
    });
}

cgbdsm_status cgbdsm_stats_get(cgbdsm_query* q, cgbdsm_stats* out) {
    return Guard([&] {
        if (!q || !out) Throw(CGBDSM_ERR_INVALID, "null argument");
        ReadCounters(q);
        q->st.dv_count = q->dv;
        q->st.pool_used = q->h_cnt->pool_top;
        q->st.pool_cap = q->pool.cap;
        q->st.last_rows_touched = q->h_cnt->rows_touched;
        *out = q->st;
    });
}

cgbdsm_status cgbdsm_sync(cgbdsm_query* q) {
    return Guard([&] {
        if (!q) Throw(CGBDSM_ERR_INVALID, "null query");
        CGD_HIP(hipStreamSynchronize(q->s));
    });
}

}  // extern "C"
