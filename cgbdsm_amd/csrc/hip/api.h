// Internal interface shared by the C-ABI translation units of the engine
// (capi.hip: Stage 2; stage1.hip: the device Stage-1 CaLiG producer): the
// context and query objects behind the opaque handles of cgbdsm.h, and the
// helpers for device memory, counters and errors.
#pragma once
#include <atomic>
#include <deque>
#include <string>
#include <vector>

#include "engine.h"
#include "kernels.h"

namespace cgd {
struct S1Query;     // device Stage-1 state of a query (stage1.hip)
struct SplitState;  // host mirror and host share of the final DFS (split.hip)
}
using namespace cgd;

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
};

struct cgbdsm_ctx {
    int device = 0;
    uint64_t budget = 0;
    std::atomic<uint64_t> used{0};  // queries of one context may run on different threads
    int num_cus = 256;
    int wall_khz = 100000;  // device wall clock (wall_clock64) rate
};

// Rows left for the final DFS after seeds and BFS levels: count, depth and
// which of the two frontier buffers holds them.
struct Frontier {
    uint64_t n = 0;
    uint32_t depth = 0, cur = 0;
};

struct cgbdsm_query {
    cgbdsm_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    cgbdsm_plan plan{};
    DevPlan hplan{};
    DevPlan* d_plan = nullptr;
    bool have_order = false;
    uint32_t qv = 0, qe = 0, ndir = 0, ebits = 5, ushift = 59, umask = 0;
    uint32_t dv = 0, dv_cap = 0, vbits = 1, rbits = 1, W = 0;
    DBuf<MRec> m_rec;
    DBuf<URec> u_rec;
    DBuf<uint32_t> pool;
    Counters* d_cnt = nullptr;
    Counters* h_cnt = nullptr;
    DBuf<uint64_t> k_ins, k_del, k_upd, k_usort, k_alt, seg_start, caps;
    DBuf<uint32_t> flags, seg_row, sc_val, sc_pos, rs_hist, u_vals, u_rows;
    DBuf<uint8_t> scan_scratch;
    DBuf<uint64_t> masks;
    DBuf<uint32_t> front[2];
    DBuf<unsigned long long> tag_hist;  // 2 * kMaxQE counters of the seed grouping
    uint64_t n_ins = 0, n_del = 0, n_upd = 0, u_rows_upper = 0;
    uint8_t edge_ok[kMaxQE] = {};
    uint64_t bfs_row_limit = 65536, frontier_cap_rows = 0;
    uint32_t donate_min = 4;
    bool dfs_shortest = true, dfs_suffix = true;
    // Final-DFS launch shape: frontier rows taken per cursor atomic, and
    // resident blocks per CU (0: as many as fit).  The autotuner's choice for
    // rows varies with the workload (1 on the scale-22 mix, 2 on the dense
    // 6-vertex scale-20 stream, 4-8 on the scale-20 mix); the default 4 is
    // the best untuned value on the mix (2 measured 8 % slower there, 8 %
    // faster on the dense stream); cgbdsm_set_dfs_tuning changes both.
    uint32_t dfs_grab = 4, dfs_bpc = 0;
    // Smallest candidate range a DFS wavefront splits off for idle ones, and
    // the cap on the per-wave stable-intersection cache (0: all LDS left).
    uint32_t dfs_range_min = 1024, dfs_cache_cap = 0;
    // Wave-cooperative list search of the final DFS: 0 off, else log2 of the
    // shortest list searched with pivots (lists of at most 64 entries are
    // always held one entry per lane when on).  Default 64.
    uint32_t dfs_wsearch = 6;
    double deadline_ms = 120000;  // safety net; cgbdsm_set_deadline_ms(0) disables
    StealItem* steal_items = nullptr;
    unsigned long long* steal_seq = nullptr;
    uint64_t steal_cap = 1ull << 20;
    uint32_t* h_abort = nullptr;  // coherent pinned host memory, mapped
    uint32_t* d_abort = nullptr;
    int phase = 0;  // 0 idle, 1 staged, 2 applied, 3 U built, 4 masks, 5 matched
    bool u_nonempty = false;
    uint32_t last_depth = 0, last_buf = 0;
    // Frontier of the current batch (valid from cgbdsm_build_frontier or the
    // first count until the next build_local_index).
    Frontier fr;
    bool fr_ready = false;
    // Recorded on this query's stream when its frontier is complete: a
    // batched launch on another query's stream waits for it.
    hipEvent_t fr_ev = nullptr;
    uint64_t last_rows = 0;
    cgbdsm_stats st{};
    hipEvent_t ev[8] = {};
    // kernel-class timing (events on this query's stream)
    bool profile = false;
    struct KPending {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<KPending> kpend;
    std::vector<hipEvent_t> ev_pool;
    cgbdsm_ktime kt[CGBDSM_K_COUNT] = {};
    uint64_t batch_idx_bytes = 0;
    struct Resident {
        DBuf<uint64_t> ins, del, upd;
        uint64_t n_ins = 0, n_del = 0, n_upd = 0;
        uint8_t edge_ok[kMaxQE] = {};
        uint32_t dv = 0;
    };
    std::deque<Resident> fifo;
    // final-DFS job table (one entry per query of a launch)
    DBuf<DfsJob> d_jobs;
    DBuf<unsigned long long> d_rowbase;
    std::vector<DfsJob> h_jobs;
    std::vector<unsigned long long> h_rowbase;
    // GPMA layout state (packed-memory array of all main-index keys)
    int layout = CGBDSM_LAYOUT_GAMMA;
    bool inited = false;
    // Device Stage-1 state when the query's batches come from
    // cgbdsm_calig_update (stage1.hip), else null.
    cgd::S1Query* s1 = nullptr;
    // CPU/GPU split of the final DFS (cgbdsm_set_cpu_split), else null.
    cgd::SplitState* split = nullptr;
    bool count_wide = false;  // the last matching's count did not fit 64 bits
    struct Pma {
        DBuf<uint64_t> slot, shadow, lo, S, keys, keys2;
        DBuf<uint32_t> cnt, ins_cnt, win_of, gpos, wins, leaf, leaf2, keep, pos, chosen;
        DBuf<PmaWin> meta;
        uint32_t nleaves = 0, height = 0;
        uint64_t n = 0;      // keys stored
        bool stale = false;  // flat lists lag the array (after a commit)
    } pma;
};

namespace cgapi {

using namespace cgd;

extern thread_local std::string g_err;
cgbdsm_status Fail(cgbdsm_status c, const std::string& m);
[[noreturn]] void Throw(cgbdsm_status c, const std::string& m);
template <class F>
cgbdsm_status Guard(F&& f) {
    try {
        f();
        g_err.clear();
        return CGBDSM_OK;
    } catch (const ApiError& e) {
        return Fail(e.code, e.what);
    } catch (const HipError& e) {
        return Fail(e.e == hipErrorOutOfMemory ? CGBDSM_ERR_OOM : CGBDSM_ERR_DEVICE, e.what);
    } catch (const std::bad_alloc&) {
        return Fail(CGBDSM_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return Fail(CGBDSM_ERR_DEVICE, e.what());
    }
}


void CheckCanStage(const cgbdsm_query* q);
void* DevAlloc(cgbdsm_ctx* ctx, size_t bytes);
void DevFree(cgbdsm_ctx* ctx, void* p, size_t bytes);
template <class T>
void Ensure(cgbdsm_query* q, DBuf<T>& b, size_t n, bool keep = false, size_t keep_n = 0) {
    if (n <= b.cap) return;
    size_t nc = std::max<size_t>({n, b.cap + b.cap / 2, 1024});
    T* np = (T*)DevAlloc(q->ctx, nc * sizeof(T));
    if (keep && b.p && keep_n)
        CGD_HIP(hipMemcpyAsync(np, b.p, std::min(keep_n, b.cap) * sizeof(T), hipMemcpyDeviceToDevice, q->s));
    if (b.p) {
        CGD_HIP(hipStreamSynchronize(q->s));
        DevFree(q->ctx, b.p, b.cap * sizeof(T));
    }
    b.p = np;
    b.cap = nc;
}

template <class T>
void Release(cgbdsm_query* q, DBuf<T>& b) {
    DevFree(q->ctx, b.p, b.cap * sizeof(T));
    b.p = nullptr;
    b.cap = 0;
}

Tables MakeTables(cgbdsm_query* q);
void ReadCounters(cgbdsm_query* q);
unsigned long long* Cnt(cgbdsm_query* q, unsigned long long Counters::*f);
void SetCounter(cgbdsm_query* q, unsigned long long Counters::*f, uint64_t v);
void GrowDV(cgbdsm_query* q, uint32_t dv);
void EnsureKeyScratch(cgbdsm_query* q, uint64_t n);
// Device Stage 1 (stage1.hip): drop the query's Stage-1 state.
void S1Detach(cgbdsm_query* q);
// Initial main index from n boundary keys already in q->k_ins (capi.hip).
void IndexInitDevice(cgbdsm_query* q, uint64_t n, uint32_t dv_count, uint32_t* card, float* avg);
// Makes the batch whose keys are already in q->k_ins / k_del / k_upd the
// staged one (capi.hip).
void StageDeviceKeys(cgbdsm_query* q, uint64_t n_ins, uint64_t n_del, uint64_t n_upd, const uint8_t* edge_ok,
                     uint32_t dv);
// CPU/GPU split of the final DFS (split.hip).  The On* hooks keep the host
// mirror in step with what the device applies; SplitPlan gives the rows the
// host counts (0: none); SplitStart copies them and the masks to the host and
// starts the host count; SplitFinish waits for it, adds its per-edge counts
// to q->h_cnt (read after the GPU's share) and moves the ratio.
void SplitFree(cgbdsm_query* q);
void SplitInvalidate(cgbdsm_query* q);
void SplitOnApply(cgbdsm_query* q);
void SplitOnUpdateIndex(cgbdsm_query* q);
void SplitOnCommit(cgbdsm_query* q);
uint64_t SplitPlan(cgbdsm_query* q, uint64_t n);
void SplitStart(cgbdsm_query* q, const uint32_t* d_rows, uint64_t k, uint64_t n, uint32_t D);
void SplitJoin(cgbdsm_query* q, bool abort);
void SplitFinish(cgbdsm_query* q, double gpu_ms);
int KBegin(cgbdsm_query* q, int cls);
void KEnd(cgbdsm_query* q, int idx);
void KAdd(cgbdsm_query* q, int cls, uint64_t bytes, uint64_t items);

}  // namespace cgapi
