// Merged matching of all query edges on gfx950 (MatchingBitAll,
// graph/graph_gpu.cu:1261-1480).
//
//   KSeed    depth-2 rows [(ei << tag_shift) | v0, v1] for every edge tag ei
//            whose batch lists exist (writeInitialResultsAllBit,
//            kernels/enumeration.cu:727-809), all tags in one pass: one
//            wavefront per 64 owners reads the shared first-pair lists once,
//            spreads their entries over lanes; scan + one atomic per 64
//            entries reserves rows.
//   KSeed1   depth-1 rows [(ei << tag_shift) | v0] instead, when the product
//            suffix starts at position 2.
//   KTag*    counting sort of the seed rows by edge tag.
//   KExtend  one BFS level, one wavefront per row, 64 candidates per step
//            (extendBFSAllBit with write_res, kernels/enumeration.cu:811-1155).
//   KDfs     the final count: persistent wavefronts pull rows (of one or
//            several queries) from a global atomic queue; each runs an
//            iterative DFS whose per-level survivor stacks live in LDS (64 u32
//            per level).  At the last level the survivors are only counted
//            (popcount of the ballot); the product suffix of the order is
//            counted in closed form, for a whole step of siblings at once
//            when its label classes have one member each (BatchSuffix).
//
// Candidate rule at depth d (global order position d, query vertex u_d):
//   list of the first backward neighbour, in its M list or (for ei's own two
//   directions) its U list, plus its U list when that query edge's index < ei;
//   support mask bit of (ei, u_d); injective w.r.t. positions 0..d-1; present
//   in every other backward neighbour's list (binary search in M then U).
#include <mutex>

#include "engine.h"
#include "kernels.h"
#include "wave.h"

// Register budget of the final DFS (occupancy against spills), a build-time
// tuning knob: -DCGD_DFS_WAVES_PER_EU=N asks for N resident waves per SIMD
// (3: 168 VGPRs with some spills, measured best; 2 and 4 are slower; 0: no cap).
#ifndef CGD_DFS_WAVES_PER_EU
#define CGD_DFS_WAVES_PER_EU 3
#endif
#if CGD_DFS_WAVES_PER_EU > 0
#define CGD_DFS_ATTR __attribute__((amdgpu_waves_per_eu(CGD_DFS_WAVES_PER_EU)))
#else
#define CGD_DFS_ATTR
#endif

namespace cgd {

namespace {

constexpr int kThreads = 256;
constexpr int kWpb = kThreads / kWave;

struct List {
    const uint32_t* base;
    const uint32_t* app;
    uint32_t bl, al;
};

// Plan facts the matching kernels read in their inner loops, kept where a
// read costs no vector-memory round trip: per-block LDS tables (filled once
// at kernel start) and uniform registers.  The plan's byte tables cannot be
// read with scalar loads, and a vector load of a plan byte ahead of every
// list access would put one more dependent trip to L2 on every step.
struct PlanLds {
    uint16_t pair[kMaxQV][kMaxQV];  // backward neighbour k of position d: dir | position << 8
    uint64_t app[kMaxQE][2];        // edge tag ei: dirs (bits 0..127) whose undirected edge index < ei
    uint32_t ones[kMaxQE];          // edge tag ei: query vertices whose support tile is all ones
    uint32_t own[kMaxQE];           // edge tag ei: its two dirs, sa | sb << 8
    uint32_t dep_mask[kMaxQV];      // positions whose candidate set has position x as backward neighbour
    uint64_t vs, cls, nbn;          // per position, 4 bits each: query vertex, label class, backward neighbours
};

__device__ void FillPlanLds(const Tables& T, PlanLds* L) {
    const cgbdsm_plan& p = T.P->p;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < (uint32_t)(kMaxQV * kMaxQV); i += blockDim.x) {
        const uint32_t d = i / kMaxQV, k = i % kMaxQV;
        uint16_t v = 0;
        if (d < p.qv && k < (uint32_t)(p.go_bni_offs[d + 1] - p.go_bni_offs[d])) {
            const uint32_t bq = p.go_bni[p.go_bni_offs[d] + k];
            v = (uint16_t)(p.eidx[p.go_vs[bq] * kMaxQV + p.go_vs[d]] | (bq << 8));
        }
        L->pair[d][k] = v;
    }
    for (uint32_t e = t; e < (uint32_t)kMaxQE; e += blockDim.x) {
        uint64_t a0 = 0, a1 = 0;
        uint32_t ones = 0, own = 0;
        if (e < p.qe) {
            for (uint32_t dir = 0; dir < p.ndir; dir++)
                if (p.dir_to_edge[dir] < e) {
                    if (dir < 64) a0 |= 1ull << dir;
                    else a1 |= 1ull << (dir - 64);
                }
            for (uint32_t q = 0; q < p.qv; q++)
                if (T.P->all_ones[e][q]) ones |= 1u << q;
            own = (uint32_t)T.P->sa[e] | ((uint32_t)T.P->sb[e] << 8);
        }
        L->app[e][0] = a0;
        L->app[e][1] = a1;
        L->ones[e] = ones;
        L->own[e] = own;
    }
    for (uint32_t x = t; x < (uint32_t)kMaxQV; x += blockDim.x) L->dep_mask[x] = T.P->dep_mask[x];
    if (t == 0) {
        uint64_t vs = 0, cls = 0, nbn = 0;
        for (uint32_t d = 0; d < p.qv; d++) {
            vs |= (uint64_t)p.go_vs[d] << (4 * d);
            cls |= (uint64_t)p.go_lcls[d] << (4 * d);
            nbn |= (uint64_t)(p.go_bni_offs[d + 1] - p.go_bni_offs[d]) << (4 * d);
        }
        L->vs = vs;
        L->cls = cls;
        L->nbn = nbn;
    }
    __syncthreads();
}

// The same tables filled by one wavefront into its own copy (the final DFS
// switches between queries per wavefront).
__device__ void FillPlanLdsWave(const Tables& T, PlanLds* L) {
    const cgbdsm_plan& p = T.P->p;
    const uint32_t t = LaneId();
    for (uint32_t i = t; i < (uint32_t)(kMaxQV * kMaxQV); i += kWave) {
        const uint32_t d = i / kMaxQV, k = i % kMaxQV;
        uint16_t v = 0;
        if (d < p.qv && k < (uint32_t)(p.go_bni_offs[d + 1] - p.go_bni_offs[d])) {
            const uint32_t bq = p.go_bni[p.go_bni_offs[d] + k];
            v = (uint16_t)(p.eidx[p.go_vs[bq] * kMaxQV + p.go_vs[d]] | (bq << 8));
        }
        L->pair[d][k] = v;
    }
    for (uint32_t e = t; e < (uint32_t)kMaxQE; e += kWave) {
        uint64_t a0 = 0, a1 = 0;
        uint32_t ones = 0, own = 0;
        if (e < p.qe) {
            for (uint32_t dir = 0; dir < p.ndir; dir++)
                if (p.dir_to_edge[dir] < e) {
                    if (dir < 64) a0 |= 1ull << dir;
                    else a1 |= 1ull << (dir - 64);
                }
            for (uint32_t q = 0; q < p.qv; q++)
                if (T.P->all_ones[e][q]) ones |= 1u << q;
            own = (uint32_t)T.P->sa[e] | ((uint32_t)T.P->sb[e] << 8);
        }
        L->app[e][0] = a0;
        L->app[e][1] = a1;
        L->ones[e] = ones;
        L->own[e] = own;
    }
    if (t < (uint32_t)kMaxQV) L->dep_mask[t] = T.P->dep_mask[t];
    if (t == 0) {
        uint64_t vs = 0, cls = 0, nbn = 0;
        for (uint32_t d = 0; d < p.qv; d++) {
            vs |= (uint64_t)p.go_vs[d] << (4 * d);
            cls |= (uint64_t)p.go_lcls[d] << (4 * d);
            nbn |= (uint64_t)(p.go_bni_offs[d + 1] - p.go_bni_offs[d]) << (4 * d);
        }
        L->vs = vs;
        L->cls = cls;
        L->nbn = nbn;
    }
    WaveSync();
}

// Uniform view of the plan for one edge tag.  The plan tables are addressed
// as LDS (a 32-bit address, not a generic pointer).
typedef const __attribute__((address_space(3))) PlanLds LdsPlan;
struct Ctx {
    LdsPlan* L;
    uint64_t vs, cls, nbn;
    uint32_t qv, ei, sa, sb, ones;
    uint64_t app0, app1;
};

__device__ __forceinline__ Ctx MakeCtx(const Tables& T, const PlanLds* L) {
    Ctx X;
    X.L = (LdsPlan*)L;
    X.vs = Uniform64(L->vs);
    X.cls = Uniform64(L->cls);
    X.nbn = Uniform64(L->nbn);
    X.qv = T.qv;
    X.ei = 0;
    X.sa = X.sb = 0xFFu;
    X.ones = 0;
    X.app0 = X.app1 = 0;
    return X;
}

__device__ __forceinline__ void SetTag(Ctx& X, uint32_t ei) {
    X.ei = ei;
    const uint32_t own = Uniform(X.L->own[ei]);
    X.sa = own & 0xFFu;
    X.sb = own >> 8;
    X.ones = Uniform(X.L->ones[ei]);
    X.app0 = Uniform64(X.L->app[ei][0]);
    X.app1 = Uniform64(X.L->app[ei][1]);
}

__device__ __forceinline__ uint32_t Vs(const Ctx& X, uint32_t d) { return (uint32_t)(X.vs >> (4 * (d & 15u))) & 15u; }
__device__ __forceinline__ uint32_t Cls(const Ctx& X, uint32_t d) { return (uint32_t)(X.cls >> (4 * (d & 15u))) & 15u; }
__device__ __forceinline__ uint32_t Nbn(const Ctx& X, uint32_t d) { return (uint32_t)(X.nbn >> (4 * (d & 15u))) & 15u; }
__device__ __forceinline__ uint32_t Pair(const Ctx& X, uint32_t d, uint32_t k) { return Uniform(X.L->pair[d][k]); }

__device__ __forceinline__ bool TileBit(const Ctx& X, const Tables& T, uint32_t qv, uint32_t v) {
    if ((X.ones >> qv) & 1u) return true;
    return (LdG(T.masks + ((uint64_t)X.ei * kMaxQV + qv) * T.W + (v >> 6)) >> (v & 63)) & 1ull;
}

// Index records read as plain words through global loads (a struct copy
// through the cast pointer can come out as flat loads once inlined).
__device__ __forceinline__ URec LdURec(const URec* p) {
    const uint64_t w = LdG((const uint64_t*)p);
    return URec{(uint32_t)w, (uint32_t)(w >> 32)};
}
__device__ __forceinline__ MRec LdMRec(const MRec* p) {
    typedef unsigned int U4 __attribute__((ext_vector_type(4)));
    const U4 w = LdG((const U4*)p);
    return MRec{(uint64_t)w.x | ((uint64_t)w.y << 32), w.z, w.w};
}

// List of directed query edge `dir` at data vertex `owner` under the
// current edge tag: the tag's own two directions read U, every other one M,
// with U appended when that direction's undirected edge precedes the tag.
__device__ __forceinline__ List Resolve(const Ctx& X, const Tables& T, uint32_t dir, uint32_t owner) {
    List L;
    const uint64_t r = (uint64_t)owner * T.ndir + dir;
    if (dir == X.sa || dir == X.sb) {
        const URec u = LdURec(T.u_rec + r);
        L.base = T.u_vals + u.off;
        L.bl = u.len;
    } else {
        const MRec m = LdMRec(T.m_rec + r);
        L.base = T.pool + m.off;
        L.bl = m.len;
    }
    const bool app = dir < 64 ? ((X.app0 >> dir) & 1ull) : ((X.app1 >> (dir - 64)) & 1ull);
    if (app) {
        const URec u = LdURec(T.u_rec + r);
        L.app = T.u_vals + u.off;
        L.al = u.len;
    } else {
        L.app = nullptr;
        L.al = 0;
    }
    return L;
}

__device__ __forceinline__ uint32_t At(const List& L, uint32_t i) { return i < L.bl ? LdG(L.base + i) : LdG(L.app + (i - L.bl)); }

// Candidate c for position d given positions 0..d-1 in `matched` (LDS,
// uniform): support-mask bit, injective, present in the list of every
// backward neighbour after the first (whose list supplied c).
__device__ __forceinline__ bool Accept(const Ctx& X, const Tables& T, uint32_t d, const uint32_t* matched, uint32_t c) {
    if (!TileBit(X, T, Vs(X, d), c)) return false;
    for (uint32_t i = 0; i < d; i++)
        if (matched[i] == c) return false;
    const uint32_t nbn = Nbn(X, d);
    for (uint32_t k = 1; k < nbn; k++) {
        const uint32_t pr = Pair(X, d, k);
        const List L = Resolve(X, T, pr & 0xFFu, Uniform(matched[pr >> 8]));
        if (!SortedContains(L.base, L.bl, c) && !(L.al && SortedContains(L.app, L.al, c))) return false;
    }
    return true;
}

// Depth-2 seeds of every active edge tag in one pass.  Position 0 -> 1 of
// the global order is one query direction `idx` for every tag, so the list
// of an owner v0 is shared: M(v0, idx) for tags other than idx's own edge
// (plus U(v0, idx) for tags after that edge), U(v0, idx) alone for the own
// edge.  A wavefront takes 64 owners (one support-mask word per tag), reads
// their list records once, and spreads the concatenated list entries over
// its lanes (an owner with a long list no longer stalls the other 63); each
// entry then emits one row per tag whose masks hold both ends.
__global__ void __launch_bounds__(kThreads) KSeed(Tables T, EiList EL, uint32_t* __restrict__ rows, uint64_t cap_rows,
                                                  Counters* __restrict__ cnt) {
    __shared__ const uint64_t* s_tile1[kMaxQE];
    __shared__ uint32_t s_tag[kMaxQE];
    __shared__ uint32_t s_incl[kWpb][kWave], s_lenm[kWpb][kWave], s_offu[kWpb][kWave];
    __shared__ uint64_t s_m0[kWpb][kWave], s_offm[kWpb][kWave];
    const cgbdsm_plan& p = T.P->p;
    const uint32_t q0 = p.go_vs[0], q1 = p.go_vs[1];
    const uint32_t idx = p.eidx[q0 * kMaxQV + q1];
    const uint32_t own = p.dir_to_edge[idx];
    const uint32_t n = EL.n;
    const uint32_t wv = threadIdx.x / kWave, lane = LaneId();
    // Per slot k of the active-tag list (bit k of the masks below).
    uint64_t ones0 = 0, ones1 = 0, use_m = 0, use_u = 0;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t ei = EL.ei[k];
        if (T.P->all_ones[ei][q0]) ones0 |= 1ull << k;
        if (T.P->all_ones[ei][q1]) ones1 |= 1ull << k;
        if (ei != own) use_m |= 1ull << k;
        if (ei >= own) use_u |= 1ull << k;
    }
    const uint32_t tag_shift = T.P->tag_shift;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        s_tile1[k] = T.masks + ((uint64_t)EL.ei[k] * kMaxQV + q1) * T.W;
        s_tag[k] = (uint32_t)EL.ei[k] << tag_shift;
    }
    __syncthreads();
    uint32_t* incl_l = s_incl[wv];
    const uint64_t nw = (uint64_t)gridDim.x * kWpb;
    for (uint64_t w = (uint64_t)blockIdx.x * kWpb + wv; w < T.W; w += nw) {
        const uint32_t v = (uint32_t)(w * kWave + lane);
        uint64_t m0 = 0;
        for (uint32_t k = 0; k < n; k++) {
            const uint64_t word =
                ((ones0 >> k) & 1ull) ? ~0ull : LdG(T.masks + ((uint64_t)EL.ei[k] * kMaxQV + q0) * T.W + w);
            m0 |= ((word >> lane) & 1ull) << k;
        }
        if (v >= T.dv) m0 = 0;
        if (!__ballot(m0 != 0)) continue;
        const uint64_t r = (uint64_t)v * T.ndir + idx;
        uint32_t lm = 0, lu = 0, ou = 0;
        uint64_t om = 0;
        if (m0 & use_m) {
            const MRec mr = LdMRec(T.m_rec + r);
            om = mr.off;
            lm = mr.len;
        }
        if (m0 & use_u) {
            const URec ur = LdURec(T.u_rec + r);
            ou = ur.off;
            lu = ur.len;
        }
        const uint32_t incl = WaveInclusiveScan(lm + lu);
        const uint32_t total = ReadLane(incl, kWave - 1);
        if (!total) continue;
        incl_l[lane] = incl;
        s_lenm[wv][lane] = lm;
        s_offu[wv][lane] = ou;
        s_m0[wv][lane] = m0;
        s_offm[wv][lane] = om;
        WaveSync();
        for (uint32_t c0 = 0; c0 < total; c0 += kWave) {
            const uint32_t j = c0 + lane;
            uint64_t ok = 0;
            uint32_t v0 = 0, nb = 0;
            if (j < total) {
                // owner lane: first lane whose inclusive count exceeds j
                uint32_t lo = 0, hi = kWave - 1;
                while (lo < hi) {
                    const uint32_t m = (lo + hi) >> 1;
                    if (incl_l[m] > j) hi = m;
                    else lo = m + 1;
                }
                const uint32_t o = j - (lo ? incl_l[lo - 1] : 0u);
                const uint32_t lmo = s_lenm[wv][lo];
                uint64_t cand;
                if (o < lmo) {
                    nb = LdG(T.pool + s_offm[wv][lo] + o);
                    cand = s_m0[wv][lo] & use_m;
                } else {
                    nb = LdG(T.u_vals + s_offu[wv][lo] + (o - lmo));
                    cand = s_m0[wv][lo] & use_u;
                }
                v0 = (uint32_t)(w * kWave + lo);
                const uint64_t bit1 = 1ull << (nb & 63u);
                for (uint64_t cm = cand; cm; cm &= cm - 1) {
                    const uint32_t k = (uint32_t)__builtin_ctzll(cm);
                    if (((ones1 >> k) & 1ull) || (s_tile1[k][nb >> 6] & bit1)) ok |= 1ull << k;
                }
            }
            const uint32_t nr = (uint32_t)__popcll(ok);
            const uint32_t incl2 = WaveInclusiveScan(nr);
            const uint32_t tot2 = ReadLane(incl2, kWave - 1);
            if (!tot2) continue;
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(&cnt->rows_out, (unsigned long long)tot2);
            base = ReadLane64(base, 0);
            if (base + tot2 > cap_rows) {
                if (lane == 0) atomicOr(&cnt->flags, kFlagFrontier);
                continue;
            }
            uint64_t pos = base + incl2 - nr;
            for (; ok; ok &= ok - 1, pos++) {
                const uint32_t k = (uint32_t)__builtin_ctzll(ok);
                rows[2 * pos] = s_tag[k] | v0;
                rows[2 * pos + 1] = nb;
            }
        }
        WaveSync();
    }
}

// Depth-1 rows [(ei << tag_shift) | v0] for every active edge tag whose
// support mask holds v0 and whose first-pair list of v0 is not empty: the
// frontier when the product suffix starts at position 2, so that the final
// DFS scans position 1 itself and counts the suffixes of v0's candidates
// together.  Same visibility as KSeed (own edge: U; later tags: M and U;
// earlier tags: M).
__global__ void __launch_bounds__(kThreads) KSeed1(Tables T, EiList EL, uint32_t* __restrict__ rows,
                                                   uint64_t cap_rows, Counters* __restrict__ cnt) {
    __shared__ uint32_t s_tag[kMaxQE];
    const cgbdsm_plan& p = T.P->p;
    const uint32_t q0 = p.go_vs[0], q1 = p.go_vs[1];
    const uint32_t idx = p.eidx[q0 * kMaxQV + q1];
    const uint32_t own = p.dir_to_edge[idx];
    const uint32_t n = EL.n;
    const uint32_t lane = LaneId();
    uint64_t ones0 = 0, use_m = 0, use_u = 0;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t ei = EL.ei[k];
        if (T.P->all_ones[ei][q0]) ones0 |= 1ull << k;
        if (ei != own) use_m |= 1ull << k;
        if (ei >= own) use_u |= 1ull << k;
    }
    const uint32_t tag_shift = T.P->tag_shift;
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) s_tag[k] = (uint32_t)EL.ei[k] << tag_shift;
    __syncthreads();
    const uint64_t nw = (uint64_t)gridDim.x * kWpb;
    for (uint64_t w = (uint64_t)blockIdx.x * kWpb + threadIdx.x / kWave; w < T.W; w += nw) {
        const uint32_t v = (uint32_t)(w * kWave + lane);
        uint64_t m0 = 0;
        for (uint32_t k = 0; k < n; k++) {
            const uint64_t word =
                ((ones0 >> k) & 1ull) ? ~0ull : LdG(T.masks + ((uint64_t)EL.ei[k] * kMaxQV + q0) * T.W + w);
            m0 |= ((word >> lane) & 1ull) << k;
        }
        if (v >= T.dv) m0 = 0;
        if (!__ballot(m0 != 0)) continue;
        const uint64_t r = (uint64_t)v * T.ndir + idx;
        uint64_t emit = 0;
        if ((m0 & use_m) && LdMRec(T.m_rec + r).len) emit |= m0 & use_m;
        if ((m0 & use_u) && LdURec(T.u_rec + r).len) emit |= m0 & use_u;
        const uint32_t nr = (uint32_t)__popcll(emit);
        const uint32_t incl = WaveInclusiveScan(nr);
        const uint32_t tot = ReadLane(incl, kWave - 1);
        if (!tot) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&cnt->rows_out, (unsigned long long)tot);
        base = ReadLane64(base, 0);
        if (base + tot > cap_rows) {
            if (lane == 0) atomicOr(&cnt->flags, kFlagFrontier);
            continue;
        }
        uint64_t pos = base + incl - nr;
        for (; emit; emit &= emit - 1, pos++) rows[pos] = s_tag[__builtin_ctzll(emit)] | v;
    }
}

// Counting sort of depth-2 rows by edge tag (KSeed emits the tags of one
// owner together; the final DFS runs faster over tag-grouped rows, whose
// waves then share one tag's support masks in L2).  Order inside a tag is
// arbitrary.  hist[0..kMaxQE) = rows per tag, hist[kMaxQE..) = cursors.
template <uint32_t W>
__global__ void __launch_bounds__(kThreads) KTagHist(const uint32_t* __restrict__ in, uint64_t n, uint32_t tag_shift,
                                                     uint64_t per_block, unsigned long long* __restrict__ hist) {
    __shared__ uint32_t h[kMaxQE];
    for (uint32_t t = threadIdx.x; t < (uint32_t)kMaxQE; t += blockDim.x) h[t] = 0;
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * per_block, hi = lo + per_block < n ? lo + per_block : n;
    for (uint64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) atomicAdd(&h[in[W * r] >> tag_shift], 1u);
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < (uint32_t)kMaxQE; t += blockDim.x)
        if (h[t]) atomicAdd(&hist[t], (unsigned long long)h[t]);
}

__global__ void __launch_bounds__(kWave) KTagScan(unsigned long long* __restrict__ hist) {
    const uint32_t lane = LaneId();
    const unsigned long long c = lane < (uint32_t)kMaxQE ? hist[lane] : 0ull;
    const unsigned long long incl = WaveInclusiveScan(c);
    if (lane < (uint32_t)kMaxQE) hist[kMaxQE + lane] = incl - c;
}

template <uint32_t W>
__global__ void __launch_bounds__(kThreads) KTagScatter(const uint32_t* __restrict__ in, uint64_t n, uint32_t tag_shift,
                                                        uint64_t per_block, unsigned long long* __restrict__ hist,
                                                        uint32_t* __restrict__ out) {
    __shared__ uint32_t h[kMaxQE];
    __shared__ unsigned long long base[kMaxQE];
    for (uint32_t t = threadIdx.x; t < (uint32_t)kMaxQE; t += blockDim.x) h[t] = 0;
    __syncthreads();
    const uint64_t lo = (uint64_t)blockIdx.x * per_block, hi = lo + per_block < n ? lo + per_block : n;
    for (uint64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) atomicAdd(&h[in[W * r] >> tag_shift], 1u);
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < (uint32_t)kMaxQE; t += blockDim.x) {
        base[t] = h[t] ? atomicAdd(&hist[kMaxQE + t], (unsigned long long)h[t]) : 0ull;
        h[t] = 0;
    }
    __syncthreads();
    for (uint64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) {
        const uint32_t w0 = in[W * r];
        const uint32_t k = w0 >> tag_shift;
        const uint64_t pos = base[k] + atomicAdd(&h[k], 1u);
        out[W * pos] = w0;
        if (W > 1) out[W * pos + 1] = in[W * r + 1];
    }
}

__global__ void __launch_bounds__(kThreads) KExtend(Tables T, const uint32_t* __restrict__ in, uint64_t n_in,
                                                    uint32_t D, uint32_t* __restrict__ out, uint64_t cap_rows,
                                                    Counters* __restrict__ cnt) {
    __shared__ uint32_t s_matched[kWpb][kMaxQV];
    __shared__ PlanLds s_plan;
    FillPlanLds(T, &s_plan);
    Ctx X = MakeCtx(T, &s_plan);
    const uint32_t wv = threadIdx.x / kWave, lane = LaneId();
    uint32_t* matched = s_matched[wv];
    const uint32_t p0 = Pair(X, D, 0);
    const uint32_t pe = p0 & 0xFFu, pq = p0 >> 8;
    const uint32_t tag_shift = T.P->tag_shift, v_mask = T.P->v_mask;
    const uint64_t nw = (uint64_t)gridDim.x * kWpb;
    for (uint64_t r = (uint64_t)blockIdx.x * kWpb + wv; r < n_in; r += nw) {
        const uint32_t w0 = Uniform(in[r * D]);
        const uint32_t ei = w0 >> tag_shift;
        if (ei != X.ei || r == (uint64_t)blockIdx.x * kWpb + wv) SetTag(X, ei);
        if (lane < D) matched[lane] = lane == 0 ? (w0 & v_mask) : in[r * D + lane];
        WaveSync();
        const List L = Resolve(X, T, pe, Uniform(matched[pq]));
        const uint32_t total = L.bl + L.al;
        for (uint32_t off = 0; off < total; off += kWave) {
            const uint32_t i = off + lane;
            uint32_t c = 0;
            bool ok = false;
            if (i < total) {
                c = At(L, i);
                ok = Accept(X, T, D, matched, c);
            }
            const uint64_t m = __ballot(ok);
            const uint32_t k = (uint32_t)__popcll(m);
            if (!k) continue;
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(&cnt->rows_out, (unsigned long long)k);
            base = __shfl(base, 0, kWave);
            if (base + k > cap_rows) {
                if (lane == 0) atomicOr(&cnt->flags, kFlagFrontier);
                continue;
            }
            if (ok) {
                uint32_t* o = out + (base + RankBelow(m)) * (uint64_t)(D + 1);
                o[0] = w0;
                for (uint32_t j = 1; j < D; j++) o[j] = matched[j];
                o[D] = c;
            }
        }
        WaveSync();
    }
}

struct DfsLevel {
    uint32_t off, cnt, pos, end;  // scan offset, survivors, next survivor, range end (exclusive)
    uint32_t sel;                 // backward neighbour whose list is scanned (kNoSel: not chosen yet)
};

// Counts saturate at kCountSat: a count that does not fit 64 bits stays
// there through sums and products (a product with an empty factor is 0),
// and the host turns it into CGBDSM_ERR_OVERFLOW instead of a wrapped value.
constexpr unsigned long long kCountSat = ~0ull;
__device__ __forceinline__ unsigned long long SatAdd(unsigned long long a, unsigned long long b) {
    const unsigned long long r = a + b;
    return r < a ? kCountSat : r;
}
__device__ __forceinline__ unsigned long long SatMul(unsigned long long a, unsigned long long b) {
    unsigned long long r;
    return __builtin_mul_overflow(a, b, &r) ? kCountSat : r;
}

constexpr uint32_t kFull = 0xFFFFFFFFu;
constexpr uint32_t kNoSel = 0xFFu;
constexpr uint32_t kNoSkip = 0xFFu;

__device__ __forceinline__ bool AbortRequested(const StealQueue& Q) {
    return __hip_atomic_load(Q.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
}

__device__ __forceinline__ unsigned long long LoadAgent(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Steal-queue items cross XCDs, whose L2s are not coherent with each other.
// Agent-scope acquire/release fences would keep them coherent by writing
// back or invalidating the whole L2 at every hand-off, which starves every
// other wave on the XCD.  Instead every item word is written and read with
// agent-scope atomic accesses (coherent without cache maintenance) and
// ordered with wait-count fences: the donor's item stores complete before
// its sequence store, and the thief reads the item only after seeing it.
__device__ __forceinline__ void PutAgent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t GetAgent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Statistics words of one final-DFS wave, in LDS: added to without waiting
// (lane 0) and read once at the end.  Held in scalar registers they were
// spilled to memory by the pressure of the candidate loops, a round trip on
// every step.
struct WaveStats {
    unsigned long long steps, step_lanes, batched, donated, scanned, probes;
};
typedef __attribute__((address_space(3))) WaveStats LdsStats;
__device__ __forceinline__ void StatAdd(__attribute__((address_space(3))) unsigned long long& f, unsigned long long v) {
    if (LaneId() == 0) __hip_atomic_fetch_add(&f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ void PublishItem(StealQueue& Q, uint64_t slot, unsigned long long ticket) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0);
    __hip_atomic_store(&Q.seq[slot], ticket + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// List of backward neighbour k of position d (owner = its matched vertex);
// k must be the same in every lane.
__device__ __forceinline__ List BnList(const Ctx& X, const Tables& T, uint32_t d, uint32_t k, const uint32_t* matched) {
    const uint32_t pr = Pair(X, d, k);
    return Resolve(X, T, pr & 0xFFu, Uniform(matched[pr >> 8]));
}

// Occurrences test by binary search in both parts of a list (probes counted).
__device__ __forceinline__ bool InList(const List& L, uint32_t c, uint32_t& probes) {
    uint32_t lo = 0, hi = L.bl;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        probes++;
        if (LdG(L.base + m) < c) lo = m + 1;
        else hi = m;
    }
    if (lo < L.bl) {
        probes++;
        if (LdG(L.base + lo) == c) return true;
    }
    if (!L.al) return false;
    lo = 0;
    hi = L.al;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        probes++;
        if (LdG(L.app + m) < c) lo = m + 1;
        else hi = m;
    }
    if (lo < L.al) {
        probes++;
        return LdG(L.app + lo) == c;
    }
    return false;
}

// Candidate c for position d: support-mask bit, distinct from positions
// 0..ninj-1, present in the list of every backward neighbour except `skip`.
__device__ __forceinline__ bool AcceptSel(const Ctx& X, const Tables& T, uint32_t d, const uint32_t* matched, uint32_t c,
                                          uint32_t skip, uint32_t ninj, uint32_t& probes) {
    if (!TileBit(X, T, Vs(X, d), c)) return false;
    for (uint32_t i = 0; i < ninj; i++)
        if (matched[i] == c) return false;
    const uint32_t nbn = Nbn(X, d);
    for (uint32_t k = 0; k < nbn; k++) {
        if (k == skip) continue;
        if (!InList(BnList(X, T, d, k, matched), c, probes)) return false;
    }
    return true;
}

// Lists at least this long are searched by the whole wavefront at once.
constexpr uint32_t kWaveSearchMin = 256;

// Membership test of every lane's candidate c (lanes with `act`) in one list
// L shared by the wavefront (every lane calls this).  A long base part is cut
// into 64 equal buckets: lane l loads the last entry of bucket l (64
// independent loads, one round trip), a candidate finds its bucket among
// those pivots with shuffles and binary-searches only inside it, which saves
// log2(64) dependent loads per candidate.  Short parts and the appended part
// are searched per lane as InList does.
__device__ __forceinline__ bool InListWave(const List& L, uint32_t c, bool act, uint32_t& probes, uint32_t wmin) {
    bool found = false;
    if (L.bl >= wmin) {
        const uint32_t lane = LaneId();
        const uint32_t piv = LdG(L.base + ((uint32_t)(((uint64_t)(lane + 1) * L.bl) >> 6) - 1));
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t step = 32; step >= 1; step >>= 1) {
            const uint32_t pv = __shfl(piv, (int)(pos + step - 1), kWave);
            if (pv < c) pos += step;
        }
        const uint32_t top = ReadLane(piv, kWave - 1);
        if (act) {
            probes++;
            if (c <= top) {
                uint32_t lo = (uint32_t)(((uint64_t)pos * L.bl) >> 6);
                uint32_t hi = (uint32_t)(((uint64_t)(pos + 1) * L.bl) >> 6);
                while (lo < hi) {
                    const uint32_t m = (lo + hi) >> 1;
                    probes++;
                    if (LdG(L.base + m) < c) lo = m + 1;
                    else hi = m;
                }
                found = lo < L.bl && LdG(L.base + lo) == c;
            }
        }
    } else if (act) {
        uint32_t lo = 0, hi = L.bl;
        while (lo < hi) {
            const uint32_t m = (lo + hi) >> 1;
            probes++;
            if (LdG(L.base + m) < c) lo = m + 1;
            else hi = m;
        }
        found = lo < L.bl && LdG(L.base + lo) == c;
    }
    if (act && !found && L.al) {
        uint32_t lo = 0, hi = L.al;
        while (lo < hi) {
            const uint32_t m = (lo + hi) >> 1;
            probes++;
            if (LdG(L.app + m) < c) lo = m + 1;
            else hi = m;
        }
        found = lo < L.al && LdG(L.app + lo) == c;
    }
    return found;
}

// InListWave, with a base part of at most 64 entries loaded whole (one
// coalesced load, one entry per lane) and searched by shuffles: the lanes'
// candidates need no dependent memory round trip at all.  Every lane calls.
__device__ __forceinline__ bool InListLanes(const List& L, uint32_t c, bool act, uint32_t& probes, uint32_t wmin) {
    if (L.bl > (uint32_t)kWave) return InListWave(L, c, act, probes, wmin);
    const uint32_t lane = LaneId();
    const uint32_t held = lane < L.bl ? LdG(L.base + lane) : ~0u;
    const uint32_t p = RankInLanes(held, c);
    const uint32_t at = __shfl(held, (int)p, kWave);
    bool found = act && p < L.bl && at == c;
    if (act) probes++;
    if (act && !found && L.al) {
        uint32_t lo = 0, hi = L.al;
        while (lo < hi) {
            const uint32_t m = (lo + hi) >> 1;
            probes++;
            if (LdG(L.app + m) < c) lo = m + 1;
            else hi = m;
        }
        found = lo < L.al && LdG(L.app + lo) == c;
    }
    return found;
}

// AcceptSel for a step of candidates, every lane calling (lanes with `act`
// hold a candidate); the backward-neighbour lists are searched with
// InListLanes while any lane still holds an accepted candidate.
__device__ __forceinline__ bool AcceptSelWave(const Ctx& X, const Tables& T, uint32_t d, const uint32_t* matched,
                                              uint32_t c, bool act, uint32_t skip, uint32_t ninj,
                                              uint32_t& probes, uint32_t wmin) {
    bool ok = act && TileBit(X, T, Vs(X, d), c);
    for (uint32_t i = 0; i < ninj; i++)
        if (matched[i] == c) ok = false;
    const uint32_t nbn = Nbn(X, d);
    for (uint32_t k = 0; k < nbn; k++) {
        if (k == skip) continue;
        if (!__ballot(ok)) break;
        ok = InListLanes(BnList(X, T, d, k, matched), c, ok, probes, wmin);
    }
    return ok;
}

// Backward neighbour of position d with the shortest visible list (lowest
// index on ties).  Scanning it instead of the first one enumerates the same
// set when no list holds a value twice.
__device__ __forceinline__ uint32_t ChooseList(const Ctx& X, const Tables& T, uint32_t d, const uint32_t* matched) {
    const uint32_t nbn = Nbn(X, d);
    if (nbn <= 1) return 0;
    const uint32_t lane = LaneId();
    uint32_t key = 0xFFFFFFFFu;
    if (lane < nbn) {
        const uint32_t pr = X.L->pair[d][lane];
        const List L = Resolve(X, T, pr & 0xFFu, matched[pr >> 8]);
        const uint32_t n = L.bl + L.al;
        key = ((n < (1u << 27) ? n : (1u << 27) - 1) << 4) | lane;
    }
    return WaveMin(key) & 15u;
}

// The cached stable intersection takes the LDS that the other per-wave
// tables leave at full occupancy (sized per launch: DfsLdsShape);
// intersections that do not fit take the uncached path.  768 -> 1536 u32 took
// the dense 6-vertex bench stream from 30.4 to 26.4 ms per DFS launch.

__device__ __forceinline__ bool LdsContains(const uint32_t* a, uint32_t n, uint32_t c) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (a[m] < c) lo = m + 1;
        else hi = m;
    }
    return lo < n && a[lo] == c;
}

// Per-wave cache of S = the lists of position stab_pos from its shallower
// backward neighbours, intersected, with its support-mask bit applied, in
// LDS as two sorted runs: buf[0..la) from the source list's main part and
// buf[la..len) from its appended update part (the parts are disjoint but
// not ordered against each other).  state 0: stale, 1: valid, 2: larger
// than the buffer (not cached).
struct StabCache {
    uint32_t* buf;
    uint32_t cap;  // u32 the buffer holds
    uint32_t len, la, state;
    uint32_t pos, vk, deps;  // the cached position, its deepest backward neighbour, positions S depends on
};

__device__ __forceinline__ bool StabContains(const StabCache& sc, uint32_t c) {
    return LdsContains(sc.buf, sc.la, c) || LdsContains(sc.buf + sc.la, sc.len - sc.la, c);
}

// Fills the stable-intersection cache for position jf (its shallower
// backward neighbours' lists intersected, support-mask bit applied): scans
// the shortest of those lists and checks the others; state 2 when the
// result does not fit the buffer.
__device__ __forceinline__ void StabBuild(const Ctx& X, const Tables& T, uint32_t jf, const uint32_t* matched,
                                          StabCache& sc, uint32_t& scanned, uint32_t& probes) {
    const uint32_t lane = LaneId();
    const uint32_t vk = sc.vk;
    const uint32_t nbn = Nbn(X, jf);
    // Build S from the shortest stable list, checking the others.
    uint32_t best = 0xFFFFFFFFu, bk = 0;
    for (uint32_t k = 0; k < nbn; k++) {
        if (k == vk) continue;
        const List L = BnList(X, T, jf, k, matched);
        const uint32_t n = Uniform(L.bl + L.al);
        if (n < best) {
            best = n;
            bk = k;
        }
    }
    const List S0 = BnList(X, T, jf, bk, matched);
    uint32_t len = 0, la = 0;
    bool big = false;
    for (uint32_t off = 0; off < best; off += kWave) {
        const uint32_t i = off + lane;
        bool ok = false;
        uint32_t c = 0;
        if (i < best) {
            scanned++;
            c = At(S0, i);
            ok = TileBit(X, T, Vs(X, jf), c);
            for (uint32_t k = 0; k < nbn && ok; k++)
                if (k != vk && k != bk) ok = InList(BnList(X, T, jf, k, matched), c, probes);
        }
        const uint64_t m = __ballot(ok);
        const uint32_t cnt = (uint32_t)__popcll(m);
        if (len + cnt > sc.cap) {
            big = true;
            break;
        }
        if (ok) sc.buf[len + RankBelow(m)] = c;
        len += cnt;
        la += (uint32_t)__popcll(__ballot(ok && i < S0.bl));
    }
    WaveSync();
    sc.len = len;
    sc.la = la;
    sc.state = big ? 2u : 1u;
}

// Size of C_B = the intersection of the candidate sets of the positions in
// B (bit mask, all at or after j0): every backward-neighbour list of every
// position of B, and the support-mask bit of each.  `raw` leaves the
// positions 0..j0-1 in; `hits` counts those of them that lie in C_B (only
// positions of the same label class can).  All lists must be sets.
__device__ __forceinline__ void SetSize(const Ctx& X, const Tables& T, uint32_t B, uint32_t j0, const uint32_t* matched, bool need_raw,
                        uint32_t& raw, uint32_t& hits, StabCache& sc, uint32_t& scanned,
                        uint32_t& probes) {
    const uint32_t lane = LaneId();
    const uint32_t jf = (uint32_t)__builtin_ctz(B);
    const uint32_t cls = Cls(X, jf);
    // Only earlier positions with the same label can lie in C_B.
    const bool may_hit = __ballot(lane < j0 && Cls(X, lane) == cls) != 0;
    if (!need_raw && !may_hit) {
        hits = 0;
        return;
    }
    if (sc.pos < kMaxQV && B == (1u << sc.pos)) {
        const uint32_t vk = sc.vk;
        if (sc.state == 0) StabBuild(X, T, jf, matched, sc, scanned, probes);
        if (sc.state == 1) {
            const List V = BnList(X, T, jf, vk, matched);
            const uint32_t nv = Uniform(V.bl + V.al);
            if (need_raw) {
                uint32_t n = 0;
                // Scan S (LDS) and binary-search V in global memory, or scan V
                // (one coalesced load per 64) and binary-search S in LDS:
                // whichever takes fewer dependent global round trips.
                const uint32_t it_s = (sc.len + kWave - 1) / kWave, it_v = (nv + kWave - 1) / kWave;
                const uint32_t depth_v = 32u - (uint32_t)__builtin_clz(nv | 1u);
                if (it_s * (depth_v + 1) < it_v) {
                    for (uint32_t off = 0; off < sc.len; off += kWave) {
                        const uint32_t i = off + lane;
                        bool ok = false;
                        if (i < sc.len) {
                            scanned++;
                            ok = InList(V, sc.buf[i], probes);
                        }
                        n += (uint32_t)__popcll(__ballot(ok));
                    }
                } else {
                    for (uint32_t off = 0; off < nv; off += kWave) {
                        const uint32_t i = off + lane;
                        bool ok = false;
                        if (i < nv) {
                            scanned++;
                            ok = StabContains(sc, At(V, i));
                        }
                        n += (uint32_t)__popcll(__ballot(ok));
                    }
                }
                raw = n;
            }
            bool hit = false;
            if (lane < j0 && Cls(X, lane) == cls && raw)
                hit = StabContains(sc, matched[lane]) && InList(V, matched[lane], probes);
            hits = (uint32_t)__popcll(__ballot(hit));
            return;
        }
    }
    // Source: the shortest list over all (position, backward neighbour) pairs.
    uint32_t best = 0xFFFFFFFFu, src_j = jf, src_k = 0, npairs = 0;
    for (uint32_t bb = B; bb; bb &= bb - 1) {
        const uint32_t j = (uint32_t)__builtin_ctz(bb);
        const uint32_t nbn = Nbn(X, j);
        for (uint32_t k = 0; k < nbn; k++, npairs++) {
            const List L = BnList(X, T, j, k, matched);
            const uint32_t n = Uniform(L.bl + L.al);
            if (n < best) {
                best = n;
                src_j = j;
                src_k = k;
            }
        }
    }
    const bool single = (B & (B - 1)) == 0 && npairs == 1 && ((X.ones >> Vs(X, jf)) & 1u);
    auto member = [&](uint32_t c, bool skip_src) -> bool {
        for (uint32_t bb = B; bb; bb &= bb - 1) {
            const uint32_t j = (uint32_t)__builtin_ctz(bb);
            if (!TileBit(X, T, Vs(X, j), c)) return false;
            const uint32_t nbn = Nbn(X, j);
            for (uint32_t k = 0; k < nbn; k++) {
                if (skip_src && j == src_j && k == src_k) continue;
                if (!InList(BnList(X, T, j, k, matched), c, probes)) return false;
            }
        }
        return true;
    };
    if (need_raw) {
        if (single || best == 0) {
            raw = best;
        } else {
            const List S = BnList(X, T, src_j, src_k, matched);
            uint32_t n = 0;
            for (uint32_t off = 0; off < best; off += kWave) {
                const uint32_t i = off + lane;
                bool ok = false;
                if (i < best) {
                    scanned++;
                    ok = member(At(S, i), true);
                }
                n += (uint32_t)__popcll(__ballot(ok));
            }
            raw = n;
        }
    }
    bool hit = false;
    if (lane < j0 && Cls(X, lane) == cls && raw) hit = member(matched[lane], false);
    hits = (uint32_t)__popcll(__ballot(hit));
}

// Number of injective assignments of the product suffix j0..qv-1 given
// positions 0..j0-1.  Positions of different label classes have disjoint
// candidate sets, so the count is the product over classes of the count
// within a class, by inclusion-exclusion over the ways members coincide:
//   1 member:  |A|
//   2 members: |A||B| - |AB|
//   3 members: |A||B||C| - |AB||C| - |AC||B| - |BC||A| + 2|ABC|
// (A_X = intersection of the members' candidate sets minus positions
// 0..j0-1; the planner keeps classes of the suffix to 3 members).  Raw
// sizes of single positions are cached per wave (valid bits in `cvalid`)
// until a position they depend on changes.
__device__ unsigned long long SuffixCount(const Ctx& X, const Tables& T, uint32_t j0, const uint32_t* matched,
                                          uint32_t* csize, uint32_t& cvalid, StabCache& sc,
                                          uint32_t& scanned, uint32_t& probes) {
    const uint32_t qv = X.qv;
    uint32_t done = 0;
    unsigned long long result = 1;
    // Classes whose sizes are all cached first: an empty one ends the count
    // before any list is read.
    // (A class is found at its first member; `seen` keeps a class skipped in
    // one pass from being found again at its later members.)
    const uint32_t cv0 = cvalid;
    for (uint32_t pass = 0; pass < 2; pass++)
    for (uint32_t j = j0, seen = done; j < qv; j++) {
        if ((seen >> j) & 1u) continue;
        uint32_t members = 0;
        for (uint32_t i = j; i < qv; i++)
            if (Cls(X, i) == Cls(X, j)) members |= 1u << i;
        seen |= members;
        if ((pass == 0) != ((members & ~cv0) == 0)) continue;
        done |= members;
        const uint32_t p0 = j;
        const uint32_t r1 = members & ~(1u << p0);
        const uint32_t p1 = r1 ? (uint32_t)__builtin_ctz(r1) : 0;
        const uint32_t r2 = r1 & ~(1u << p1);
        const uint32_t p2 = r2 ? (uint32_t)__builtin_ctz(r2) : 0;
        const uint32_t m = (uint32_t)__popc(members);
        long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // indexed by local subset
        bool empty = false;
        // Singletons first (subsets 1, 2, 4), then the joint ones.
        for (uint32_t t = 0; t < (1u << m) - 1 && !empty; t++) {
            const uint32_t S = t < m ? (1u << t) : (t == m ? 3u : t == m + 1 ? 5u : t == m + 2 ? 6u : 7u);
            const uint32_t B = ((S & 1u) ? 1u << p0 : 0u) | ((S & 2u) ? 1u << p1 : 0u) | ((S & 4u) ? 1u << p2 : 0u);
            const bool single = t < m;
            const uint32_t pi = t == 0 ? p0 : t == 1 ? p1 : p2;
            const bool cached = single && ((cvalid >> pi) & 1u);
            uint32_t raw = cached ? Uniform(csize[pi]) : 0, hits = 0;
            SetSize(X, T, B, j0, matched, !cached, raw, hits, sc, scanned, probes);
            if (single && !cached) {
                if (LaneId() == 0) csize[pi] = raw;
                cvalid |= 1u << pi;
            }
            const long long val = (long long)raw - (long long)hits;
            v[S] = val;
            if (single && val <= 0) empty = true;
        }
        if (empty) return 0;
        __int128 nc;
        if (m == 1) nc = v[1];
        else if (m == 2) nc = (__int128)v[1] * v[2] - v[3];
        else
            nc = (__int128)v[1] * v[2] * v[4] - (__int128)v[3] * v[4] - (__int128)v[5] * v[2] - (__int128)v[6] * v[1] +
                 2 * (__int128)v[7];
        if (nc <= 0) return 0;
        result = SatMul(result, (nc >> 64) ? kCountSat : (unsigned long long)nc);
    }
    return result;
}

// Per-wave LDS of BatchSuffix: per sibling lane, its varying list and the
// running candidate count.
struct BxLds {
    const uint32_t* base[kWave];
    const uint32_t* app[kWave];
    uint32_t bl[kWave], al[kWave], incl[kWave], cnt[kWave];  // al bit 31: scan the source, probe the list
    uint32_t own[kWave];  // chunk position -> sibling whose segment starts there
};

// Product-suffix counts of all survivors of one 64-candidate step at the
// last position p before the suffix (lane s holds sibling x_s when bit s of
// m is set), summed.  Every label class of the suffix has one member, so a
// sibling's count is the product over suffix positions j of |C_j| - hits_j.
// Positions whose backward neighbours lie before p have sibling-independent
// sizes (cached per wave while those positions stay); the others are
// counted for all siblings at once: their lists N_{p->j}(x_s) are laid end
// to end over the 64 lanes, each entry tested against j's other lists (or
// the cached stable intersection) and its support mask, and counted per
// sibling in LDS.  Equal to SuffixCount per sibling (identical counts).
__device__ unsigned long long BatchSuffix(const Ctx& X, const Tables& T, uint32_t p, const uint32_t* matched,
                                          uint64_t m, uint32_t x, uint32_t* csize, uint32_t& cvalid, StabCache& sc,
                                          BxLds& B, uint32_t& scanned, uint32_t& probes, uint32_t wmin) {
    const uint32_t lane = LaneId();
    const uint32_t qv = X.qv;
    const uint32_t pcls = Cls(X, p);
    unsigned long long prod = ((m >> lane) & 1ull) ? 1ull : 0ull;
    // Cheap positions first (sizes independent of the sibling, or a single
    // list with no mask to apply): a sibling they leave empty adds no
    // entries to the costly intersections that follow.
    // Three tiers: sibling-independent or a bare single list; a single list
    // filtered by its mask; intersections of several lists.
    uint32_t tier[3] = {0, 0, 0};
    for (uint32_t j = p + 1; j < qv; j++) {
        const uint32_t nbn = Nbn(X, j);
        bool varying = false;
        for (uint32_t k = 0; k < nbn; k++) varying |= (Pair(X, j, k) >> 8) == p;
        const uint32_t t = !varying || (nbn == 1 && ((X.ones >> Vs(X, j)) & 1u)) ? 0u : nbn == 1 ? 1u : 2u;
        tier[t] |= 1u << j;
    }
    for (uint32_t pass = 0; pass < 3; pass++)
    for (uint32_t jm = tier[pass]; jm; jm &= jm - 1) {
        const uint32_t j = (uint32_t)__builtin_ctz(jm);
        if (!__ballot(prod != 0)) break;
        const uint32_t nbn = Nbn(X, j);
        uint32_t kp = kNoSkip;
        for (uint32_t k = 0; k < nbn; k++)
            if ((Pair(X, j, k) >> 8) == p) kp = k;
        const uint32_t cls = Cls(X, j), qj = Vs(X, j);
        const bool same = pcls == cls;  // x_s itself may fall in C_j
        if (kp == kNoSkip) {
            // Size and hits are both sibling-independent; the hits (earlier
            // same-class positions inside C_j) hold while positions 0..p-1
            // stay (bits 16.. of cvalid, dropped at every descent above p),
            // so a level's steps after its first read both from LDS.
            const bool cached = (cvalid >> j) & 1u;
            const bool hcached = cached && ((cvalid >> (16 + j)) & 1u);
            uint32_t raw = cached ? Uniform(csize[j]) : 0u, hits = 0;
            if (hcached) {
                hits = Uniform(csize[kMaxQV + j]);
            } else {
                SetSize(X, T, 1u << j, p, matched, !cached, raw, hits, sc, scanned, probes);
                if (lane == 0) {
                    csize[j] = raw;
                    csize[kMaxQV + j] = hits;
                }
                cvalid |= (1u << j) | (1u << (16 + j));
            }
            long long val = (long long)raw - (long long)hits;
            // Is the sibling itself in C_j?  C_j's lists do not depend on
            // the sibling, so with the wave search on, the siblings (sorted,
            // one per lane) are searched in each list by the whole wavefront.
            const bool act = same && prod && val > 0;
            if (wmin && __ballot(act)) {
                bool in = act && TileBit(X, T, qj, x);
                for (uint32_t k = 0; k < nbn; k++) {
                    if (!__ballot(in)) break;
                    in = InListLanes(BnList(X, T, j, k, matched), x, in, probes, wmin);
                }
                val -= in ? 1 : 0;
            } else if (!wmin && act) {
                bool in = TileBit(X, T, qj, x);
                for (uint32_t k = 0; k < nbn && in; k++) in = InList(BnList(X, T, j, k, matched), x, probes);
                val -= in ? 1 : 0;
            }
            prod = val > 0 ? SatMul(prod, (unsigned long long)val) : 0ull;
            continue;
        }
        // j's list from p, per sibling
        const uint32_t dir = Pair(X, j, kp) & 0xFFu;
        List Ls{nullptr, nullptr, 0, 0};
        if (prod) Ls = Resolve(X, T, dir, x);
        const uint32_t len = Ls.bl + Ls.al;
        const bool ones = (X.ones >> qj) & 1u;
        uint32_t raw;
        bool use_sc = false;  // S = j's stable lists and mask, cached in LDS
        if (nbn == 1 && ones) {
            raw = len;
        } else {
            if (nbn > 1 && sc.pos == j && sc.vk == kp) {
                if (sc.state == 0) StabBuild(X, T, j, matched, sc, scanned, probes);
                use_sc = sc.state == 1;
            }
            // Sibling-independent source: the cached intersection S, or j's
            // shortest list from a position before p.
            uint32_t kmin = kNoSkip, nmin = 0xFFFFFFFFu;
            if (use_sc) {
                nmin = sc.len;
            } else {
                for (uint32_t k = 0; k < nbn; k++) {
                    if (k == kp) continue;
                    const List L = BnList(X, T, j, k, matched);
                    const uint32_t n = Uniform(L.bl + L.al);
                    if (n < nmin) {
                        nmin = n;
                        kmin = k;
                    }
                }
            }
            const List Lmin = kmin != kNoSkip ? BnList(X, T, j, kmin, matched) : List{nullptr, nullptr, 0, 0};
            // Per sibling, scan the cheaper side: the source (probing the
            // sibling's list by binary search) or the sibling's list.  Lanes
            // probing one sibling's list share the first steps of their
            // searches (same midpoints, same cache lines), so a search costs
            // far less than depth+1 separate loads: scanning the cached source
            // measured best up to 4x past the naive equal-cost point (dense
            // 6-vertex bench stream: 32.4 -> 30.4 ms per launch).
            bool from_src = false;
            if (prod && nmin != 0xFFFFFFFFu) {
                const uint32_t depth_v = 32u - (uint32_t)__builtin_clz(len | 1u);
                from_src = use_sc ? (uint64_t)nmin * (depth_v + 1) < 4ull * len : nmin < len;
            }
            const uint32_t ne = prod ? (from_src ? nmin : len) : 0u;
            const uint32_t incl = WaveInclusiveScan(ne);
            const uint32_t total = ReadLane(incl, kWave - 1);
            B.base[lane] = Ls.base;
            B.app[lane] = Ls.app;
            B.bl[lane] = Ls.bl;
            B.al[lane] = from_src ? (Ls.al | 0x80000000u) : Ls.al;
            B.incl[lane] = incl;
            B.cnt[lane] = 0;
            B.own[lane] = 0;
            WaveSync();
            const uint32_t st = incl - ne;  // this sibling's first element
            for (uint32_t c0 = 0; c0 < total; c0 += kWave) {
                const uint32_t e = c0 + lane;
                // Owner of element e: each non-empty segment marks its first
                // element in this chunk (the one covering c0 marks lane 0), and
                // a prefix max carries the marks forward.  Marks left from
                // earlier chunks of this position name earlier siblings than
                // the one covering c0, so they never win.
                if (ne && st < c0 + kWave && incl > c0) B.own[st > c0 ? st - c0 : 0u] = lane;
                WaveSync();
                uint32_t lo = B.own[lane];
                lo = max(lo, DppMove<kDppShr1>(lo));
                lo = max(lo, DppMove<kDppShr2>(lo));
                lo = max(lo, DppMove<kDppShr4>(lo));
                lo = max(lo, DppMove<kDppShr8>(lo));
                lo = max(lo, DppMove<kDppBcast15, 0xa>(lo));
                lo = max(lo, DppMove<kDppBcast31, 0xc>(lo));
                if (e < total) {
                    const uint32_t o = e - (lo ? B.incl[lo - 1] : 0u);
                    const uint32_t alw = B.al[lo];
                    scanned++;
                    bool ok;
                    if (alw & 0x80000000u) {
                        const uint32_t c = use_sc ? sc.buf[o] : At(Lmin, o);
                        const List Lo{B.base[lo], B.app[lo], B.bl[lo], alw & 0x7FFFFFFFu};
                        ok = InList(Lo, c, probes);
                        if (ok && !use_sc) {
                            ok = TileBit(X, T, qj, c);
                            for (uint32_t k = 0; k < nbn && ok; k++)
                                if (k != kp && k != kmin) ok = InList(BnList(X, T, j, k, matched), c, probes);
                        }
                    } else {
                        const uint32_t c = o < B.bl[lo] ? LdG(B.base[lo] + o) : LdG(B.app[lo] + (o - B.bl[lo]));
                        if (use_sc) {
                            ok = StabContains(sc, c);
                        } else {
                            ok = TileBit(X, T, qj, c);
                            for (uint32_t k = 0; k < nbn && ok; k++)
                                if (k != kp) ok = InList(BnList(X, T, j, k, matched), c, probes);
                        }
                    }
                    if (ok) atomicAdd(&B.cnt[lo], 1u);
                }
            }
            WaveSync();
            raw = B.cnt[lane];
        }
        // Positions before the suffix with j's label class that lie in C_j.
        uint32_t hits = 0;
        if (prod && raw) {
            // The sibling-independent part first (from the cached S when
            // valid), the sibling's own list last.
            auto member = [&](uint32_t y) -> bool {
                if (use_sc) {
                    if (!StabContains(sc, y)) return false;
                } else {
                    if (!TileBit(X, T, qj, y)) return false;
                    for (uint32_t k = 0; k < nbn; k++)
                        if (k != kp && !InList(BnList(X, T, j, k, matched), y, probes)) return false;
                }
                return InList(Ls, y, probes);
            };
            for (uint32_t i = 0; i < p; i++)
                if (Cls(X, i) == cls && member(matched[i])) hits++;
            if (same && member(x)) hits++;
        }
        const long long val = (long long)raw - (long long)hits;
        prod = val > 0 ? SatMul(prod, (unsigned long long)val) : 0ull;
    }
    // Saturating sum over the lanes (a DPP scan; lane 63 holds the total).
    unsigned long long sum = prod;
    sum = SatAdd(sum, DppMove64<kDppShr1>(sum));
    sum = SatAdd(sum, DppMove64<kDppShr2>(sum));
    sum = SatAdd(sum, DppMove64<kDppShr4>(sum));
    sum = SatAdd(sum, DppMove64<kDppShr8>(sum));
    sum = SatAdd(sum, DppMove64<kDppBcast15, 0xa>(sum));
    sum = SatAdd(sum, DppMove64<kDppBcast31, 0xc>(sum));
    return ReadLane64(sum, kWave - 1);
}

// Final DFS.  Work items are frontier rows (taken `grab` at a time from a
// global cursor) and, once those run out, partial matches stolen from the
// global queue.  A wavefront that is descending at a shallow level while
// others idle donates the tail of its survivor list (items one level
// deeper) or the rest of a long candidate range (items at the same level,
// same scanned list); `pending` counts items not yet finished, so idle
// wavefronts leave once it reaches zero.
//
// Count mode refinements (identical counts):
//  * each level scans its shortest backward-neighbour list when the batch's
//    update lists hold no duplicate value (then every list is a set);
//  * the product suffix of the order (positions cp_start..qv-1, mutually
//    non-adjacent) is never enumerated: a partial match of the positions
//    before it adds SuffixCount.
template <bool kTuples>
__global__ void __launch_bounds__(kThreads) CGD_DFS_ATTR KDfs(const DfsJob* __restrict__ jobs, uint32_t njobs,
                                                              const unsigned long long* __restrict__ row_base,
                                                              uint32_t grab, Counters* __restrict__ cnt, StealQueue Q,
                                                              uint32_t* __restrict__ tuples, uint64_t tup_cap,
                                                              uint32_t surv_levels, uint32_t stab_cap) {
    // Per-edge counts gather in LDS per wave and reach the query's counters
    // when the wave changes query or ends: a device-scope atomic per item on
    // one address would serialize.
    __shared__ unsigned long long s_pe[kWpb][kMaxQE];
    __shared__ uint32_t s_matched[kWpb][kMaxQV];
    __shared__ DfsLevel s_lvl[kWpb][kMaxQV];
    __shared__ PlanLds s_plan[kWpb];
    __shared__ uint32_t s_csize[kWpb][2 * kMaxQV];  // cached raw sizes, then cached hits
    __shared__ BxLds s_bx[kWpb];
    __shared__ WaveStats s_stats[kWpb];
    // Dynamic LDS, per wave: survivor stacks of surv_levels levels (64 u32
    // each; every job of the launch has qv - 1 <= surv_levels) and the
    // stable-intersection cache of stab_cap u32.
    extern __shared__ uint32_t s_dyn[];
    // The wave's index, read first lane: the per-wave LDS pointers below are
    // then scalar values, not per-lane copies.
    const uint32_t wv = Uniform(threadIdx.x / kWave), lane = LaneId();
    uint32_t* surv = s_dyn + wv * (surv_levels * kWave + stab_cap);
    uint32_t* stab = surv + surv_levels * kWave;
    uint32_t* matched = s_matched[wv];
    DfsLevel* lvl = s_lvl[wv];
    unsigned long long* pe = s_pe[wv];
    for (uint32_t t = lane; t < (uint32_t)kMaxQE; t += kWave) pe[t] = 0;
    const uint64_t n_rows = row_base[njobs];
    // The current query of this wavefront and what depends on it.
    uint32_t job = 0xFFFFFFFFu, D = 0, qv = 0, last = 0, tag_shift = 0, v_mask = 0, cps = 0;
    Tables T = jobs[0].T;
    const uint32_t* rows = nullptr;
    Counters* jc = nullptr;
    Ctx X = MakeCtx(T, &s_plan[wv]);
    StabCache sc{stab, stab_cap, 0, 0, 0, 0xFFu, 0, 0};
    bool sets = false, cart = false, bsfx = false, wsearch = false;
    uint32_t wmin = kWaveSearchMin;
    auto flush_pe = [&]() {
        if (!jc) return;
        for (uint32_t t = lane; t < (uint32_t)kMaxQE; t += kWave)
            if (pe[t]) {
                const unsigned long long old = atomicAdd(&jc->per_edge[t], pe[t]);
                if (pe[t] == kCountSat || old + pe[t] < old) atomicOr(&jc->count_ovf, 1ull);
                pe[t] = 0;
            }
        WaveSync();
    };
    auto use_job = [&](uint32_t j) {
        if (j == job) return;
        flush_pe();
        job = j;
        T = jobs[j].T;
        rows = jobs[j].rows;
        jc = jobs[j].cnt;
        D = jobs[j].D;
        FillPlanLdsWave(T, &s_plan[wv]);
        X = MakeCtx(T, &s_plan[wv]);
        qv = X.qv;
        last = qv - 1;
        tag_shift = T.P->tag_shift;
        v_mask = T.P->v_mask;
        sc = StabCache{stab, stab_cap, 0, 0, 0, T.P->stab_pos, T.P->stab_var_k, T.P->stab_deps};
        // Every list is a set when the update lists hold no duplicate value.
        const bool lists_are_sets = Uniform64(jc->u_dups) == 0;
        const uint32_t opts = jobs[j].opts;
        sets = (opts & kJobShortest) && lists_are_sets;
        cps = T.P->cp_start;
        cart = !kTuples && lists_are_sets && (opts & kJobSuffix) && cps < qv;
        bsfx = cart && T.P->sfx_batch;
        wsearch = (opts & kJobWaveSearch) != 0;
        // At least 64 entries: every lane's bucket pivot index is >= 0.
        wmin = (opts >> kJobWaveShift) & 31u ? 1u << ((opts >> kJobWaveShift) & 31u) : kWaveSearchMin;
        wmin = wmin < (uint32_t)kWave ? (uint32_t)kWave : wmin;
    };
    // Valid cached singleton sizes (bit per position) and, from bit 16, valid
    // cached hits of the batched suffix (dropped whenever a position changes).
    uint32_t cvalid = 0;
    // Candidates scanned and binary-search probes: 32-bit per lane (two
    // fewer registers each), folded into the wave's 64-bit totals at every
    // wall-clock poll (a lane cannot reach 2^32 within one poll period) and
    // at exit.
    uint32_t my_scanned = 0, my_probes = 0;
    LdsStats* ws = (LdsStats*)&s_stats[wv];
    if (lane == 0) ws->steps = ws->step_lanes = ws->batched = ws->donated = ws->scanned = ws->probes = 0;
    WaveSync();
    auto fold = [&]() {
        const unsigned long long a = ReadLane64(WaveInclusiveScan((unsigned long long)my_scanned), kWave - 1);
        const unsigned long long b = ReadLane64(WaveInclusiveScan((unsigned long long)my_probes), kWave - 1);
        StatAdd(ws->scanned, a);
        StatAdd(ws->probes, b);
        my_scanned = 0;
        my_probes = 0;
    };
    uint64_t r_next = 0, r_end = 0, r_first = 0, ticket = 0;
    bool primary = true, idle = false, stop = false, have_ticket = false, from_row = false;
    uint32_t iter = 0, backoff = 1;
    // Shared words (idle count, flags) and the host flag are read at most
    // once per poll_ticks / host_ticks of wall clock per wave: thousands of
    // waves polling one address would otherwise saturate the memory-side
    // path that agent-scope loads take and stall the waves that work.
    long long t_poll = Uniform64(wall_clock64()), t_host = t_poll;
    const long long t_born = t_poll;
    long long t_idle0 = 0, idle_sum = 0;
    unsigned long long idl_hint = 0;
    // Returns true when the kernel is asked to stop.
    auto poll = [&]() -> bool {
        const long long now = (long long)Uniform64((unsigned long long)wall_clock64());
        if (now - t_poll < Q.poll_ticks) return false;
        t_poll = now;
        if (__ballot(my_scanned >= (1u << 30) || my_probes >= (1u << 30))) fold();
        const bool host = now - t_host >= Q.host_ticks;
        if (host) t_host = now;
        unsigned long long idl = 0, fl = 0;
        int ab = 0;
        if (lane == 0) {
            // Waves waiting for work hold tickets at or above the tail.
            const unsigned long long h = LoadAgent(&cnt->q_head.v), t = LoadAgent(&cnt->q_tail.v);
            idl = h > t ? h - t : 0;
            fl = LoadAgent(&cnt->flags);
            if (host && AbortRequested(Q)) {
                atomicOr(&cnt->flags, kFlagAborted);
                ab = 1;
            }
        }
        // Uniform values (read first lane): branches on them stay scalar, so
        // the loops' carried state can stay in scalar registers.
        idl_hint = ReadLane64(idl, 0);
        fl = ReadLane64(fl, 0);
        return ReadLane((uint32_t)ab, 0) || (fl & kFlagAborted);
    };
    while (!stop) {
        // ---------------- obtain a work item
        uint32_t D0 = 0, ei = 0, lo0 = 0, hi0 = kFull, sel0 = kNoSel;
        bool got = false;
        if (primary) {
            if (r_next >= r_end) {
                unsigned long long r0 = 0;
                if (lane == 0) r0 = atomicAdd(&cnt->q_rows.v, (unsigned long long)grab);
                r0 = ReadLane64(r0, 0);
                r_next = r0;
                r_end = r0 + grab < n_rows ? r0 + grab : n_rows;
                r_first = r0;
                if (r0 >= n_rows) primary = false;
            }
            if (primary) {
                const uint64_t r = r_next++;
                from_row = true;
                // the query this row belongs to
                uint32_t jlo = 0, jhi = njobs;
                while (jhi - jlo > 1) {
                    const uint32_t m = (jlo + jhi) >> 1;
                    if (row_base[m] <= r) jlo = m;
                    else jhi = m;
                }
                use_job(jlo);
                const uint64_t lr = r - row_base[jlo];
                const uint32_t w0 = Uniform(LdG(rows + lr * D));
                ei = w0 >> tag_shift;
                const uint32_t mv = lane == 0 ? (w0 & v_mask) : (lane < D ? LdG(rows + lr * D + lane) : 0u);
                // A row whose tag or vertex ids are out of range (frontier
                // written out of order with this launch) is not followed into
                // the index: it is dropped and flagged, and the call fails.
                if (ei >= T.qe || __ballot(lane < D && mv >= T.dv)) {
                    if (lane == 0) {
                        atomicOr(&cnt->flags, kFlagBadRow);
                        if (r_next >= r_end) atomicSub(&cnt->q_pending.v, (unsigned long long)(r_end - r_first));
                    }
                    continue;
                }
                if (lane < D) matched[lane] = mv;
                D0 = D;
                got = true;
            }
        }
        if (!got) {
            // One ticket per idle wave (fetch-and-add, no retries): the wave
            // then waits on its own slot until a donor fills it or no work
            // is left anywhere (`pending` = rows and items not yet finished;
            // donations only come from waves holding one, so none can
            // arrive once it is zero).
            if (!have_ticket) {
                unsigned long long tk = 0;
                if (lane == 0) {
                    tk = atomicAdd(&cnt->q_head.v, 1ull);
                }
                ticket = ReadLane64(tk, 0);
                have_ticket = true;
                idle = true;
                t_idle0 = (long long)Uniform64((unsigned long long)wall_clock64());
            }
            const uint64_t slot = ticket & (Q.cap - 1);
            unsigned long long sq = 0;
            if (lane == 0) sq = LoadAgent(&Q.seq[slot]);
            sq = ReadLane64(sq, 0);
            if (sq == ticket + 1) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                StealItem* it = &Q.items[slot];
                D0 = Uniform(GetAgent(&it->depth));
                const uint32_t tag = Uniform(GetAgent(&it->ei));
                ei = tag & 0xFFu;
                sel0 = (tag >> 8) & 0xFFu;
                use_job(tag >> 16);
                lo0 = Uniform(GetAgent(&it->lo));
                hi0 = Uniform(GetAgent(&it->hi));
                if (lane < D0) matched[lane] = GetAgent(&it->v[lane]);
                got = true;
                from_row = false;
                have_ticket = false;
                backoff = 1;
                idle = false;
                idle_sum += (long long)Uniform64((unsigned long long)wall_clock64()) - t_idle0;
            } else {
                unsigned long long pend = 0;
                if (lane == 0) pend = LoadAgent(&cnt->q_pending.v);
                pend = ReadLane64(pend, 0);
                if (pend == 0) break;
                if (poll()) break;
                for (uint32_t b = 0; b < backoff; b++) __builtin_amdgcn_s_sleep(63);
                backoff = backoff < 16 ? backoff * 2 : 16;
                continue;
            }
        }
        // ---------------- DFS from position D0
        if (ei != X.ei || X.sa == 0xFFu) SetTag(X, ei);
        if (lane == 0) lvl[D0] = DfsLevel{lo0, 0, 0, hi0, sel0};
        cvalid = 0;
        sc.state = 0;
        WaveSync();
        unsigned long long count = 0;
        uint32_t d = Uniform(D0);
        for (;;) {
            // (every iteration at the batched level, whose steps are long:
            // donors must see idle wavefronts early)
            if (((++iter & 31u) == 0 || (bsfx && d + 1 == cps)) && poll()) {
                stop = true;
                break;
            }
            // The product suffix is counted, not enumerated.
            if (cart && d >= cps) {
                count = SatAdd(count, SuffixCount(X, T, d, matched, s_csize[wv], cvalid, sc, my_scanned, my_probes));
                if (d == D0) break;
                d--;
                continue;
            }
            // Positions below `term` descend into their survivors.
            const uint32_t term = cart ? cps : last;
            if (d < term) {
                const uint32_t pos = Uniform(lvl[d].pos);
                uint32_t ccount = Uniform(lvl[d].cnt);
                if (pos < ccount) {
                    // Donate the tail of this level when others are idle.
                    // Items handed out hold at least one scanned level (or, in
                    // product mode, one suffix count) below them.
                    if (Q.donate_min && idl_hint > 0 && d + (cart ? 1u : 2u) <= term &&
                        ccount - pos >= Q.donate_min) {
                        unsigned long long h = 0, t = 0;
                        if (lane == 0) {
                            h = LoadAgent(&cnt->q_head.v);
                            t = LoadAgent(&cnt->q_tail.v);
                        }
                        h = ReadLane64(h, 0);
                        t = ReadLane64(t, 0);
                        uint32_t k = ccount - pos - 1;
                        const uint32_t want = (uint32_t)(idl_hint < 8 ? 2 * idl_hint : 16);
                        k = k < want ? k : want;
                        if (k && (long long)(t - h) + (long long)k <= (long long)(Q.cap / 2)) {
                            unsigned long long t0 = 0;
                            if (lane == 0) {
                                atomicAdd(&cnt->q_pending.v, (unsigned long long)k);
                                t0 = atomicAdd(&cnt->q_tail.v, (unsigned long long)k);
                            }
                            t0 = ReadLane64(t0, 0);
                            if (lane < k) {
                                const uint64_t slot = (t0 + lane) & (Q.cap - 1);
                                StealItem* it = &Q.items[slot];
                                PutAgent(&it->depth, d + 1);
                                PutAgent(&it->ei, ei | (kNoSel << 8) | (job << 16));
                                for (uint32_t f = 0; f < d; f++) PutAgent(&it->v[f], matched[f]);
                                PutAgent(&it->v[d], surv[d * kWave + ccount - k + lane]);
                                PutAgent(&it->lo, 0);
                                PutAgent(&it->hi, kFull);
                                PublishItem(Q, slot, t0 + lane);
                            }
                            StatAdd(ws->donated, k);
                            ccount -= k;
                            idl_hint = idl_hint > k ? idl_hint - k : 0;
                            if (lane == 0) lvl[d].cnt = ccount;
                            WaveSync();
                        }
                    }
                    const uint32_t x = Uniform(surv[d * kWave + pos]);
                    if (lane == 0) {
                        lvl[d].pos = pos + 1;
                        matched[d] = x;
                        lvl[d + 1] = DfsLevel{0, 0, 0, kFull, kNoSel};
                    }
                    cvalid &= ~Uniform(s_plan[wv].dep_mask[d]) & 0xFFFFu;
                    if ((sc.deps >> d) & 1u) sc.state = 0;
                    WaveSync();
                    d++;
                    continue;
                }
            }
            // Scan the next 64 candidates of position d.
            uint32_t sel = Uniform(lvl[d].sel);
            if (sel == kNoSel) {
                sel = sets ? ChooseList(X, T, d, matched) : 0;
                if (lane == 0) lvl[d].sel = sel;
            }
            const List L = BnList(X, T, d, sel, matched);
            const uint32_t total = Uniform(L.bl + L.al);
            const uint32_t off = Uniform(lvl[d].off);
            uint32_t limit = Uniform(lvl[d].end);
            limit = limit < total ? limit : total;
            if (off >= limit) {
                if (d == D0) break;
                d--;
                continue;
            }
            // Split the rest of this level's candidate range with idle wavefronts.
            // At the position before a batched product suffix every candidate
            // carries a whole suffix count and its survivors are not descended
            // into, so its ranges are split down to one step each (smaller
            // pieces would repeat the per-item set-up of the suffix sizes).
            const bool blev = bsfx && d + 1 == cps;
            const uint32_t rmin = blev ? kWave : Q.range_min;
            if (Q.donate_min && idl_hint > 0 && limit - off >= 2 * rmin) {
                unsigned long long h = 0, t = 0;
                if (lane == 0) {
                    h = LoadAgent(&cnt->q_head.v);
                    t = LoadAgent(&cnt->q_tail.v);
                }
                h = ReadLane64(h, 0);
                t = ReadLane64(t, 0);
                const uint32_t rem = limit - off;
                uint32_t parts = (uint32_t)(idl_hint < 15 ? idl_hint + 1 : 16);
                if (parts > rem / rmin) parts = rem / rmin;
                if (parts >= 2 && (long long)(t - h) + (long long)parts <= (long long)(Q.cap / 2)) {
                    const uint32_t piece = ((rem + parts - 1) / parts + kWave - 1) & ~(uint32_t)(kWave - 1);
                    const uint32_t k = (rem - 1) / piece;  // donated pieces after the kept one
                    unsigned long long t0 = 0;
                    if (lane == 0) {
                        atomicAdd(&cnt->q_pending.v, (unsigned long long)k);
                        t0 = atomicAdd(&cnt->q_tail.v, (unsigned long long)k);
                    }
                    t0 = ReadLane64(t0, 0);
                    if (lane < k) {
                        const uint64_t slot = (t0 + lane) & (Q.cap - 1);
                        StealItem* it = &Q.items[slot];
                        PutAgent(&it->depth, d);
                        PutAgent(&it->ei, ei | (sel << 8) | (job << 16));
                        for (uint32_t f = 0; f < d; f++) PutAgent(&it->v[f], matched[f]);
                        const uint32_t plo = off + (lane + 1) * piece;
                        PutAgent(&it->lo, plo);
                        PutAgent(&it->hi, plo + piece < limit ? plo + piece : limit);
                        PublishItem(Q, slot, t0 + lane);
                    }
                    StatAdd(ws->donated, k);
                    idl_hint = idl_hint > k ? idl_hint - k : 0;
                    limit = off + piece;
                    if (lane == 0) lvl[d].end = limit;
                    WaveSync();
                }
            }
            const uint32_t i = off + lane;
            uint32_t c = 0;
            bool ok = false;
            StatAdd(ws->steps, 1);
            StatAdd(ws->step_lanes, limit - off < kWave ? limit - off : kWave);
            if (wsearch) {
                if (i < limit) {
                    c = At(L, i);
                    my_scanned++;
                }
                ok = AcceptSelWave(X, T, d, matched, c, i < limit, sel, d, my_probes, wmin);
            } else if (i < limit) {
                c = At(L, i);
                my_scanned++;
                ok = AcceptSel(X, T, d, matched, c, sel, d, my_probes);
            }
            const uint64_t m = __ballot(ok);
            // Last position before the product suffix: count the suffix of
            // every survivor of this step at once instead of descending.
            if (blev) {
                if (m) {
                    count = SatAdd(count, BatchSuffix(X, T, d, matched, m, c, s_csize[wv], cvalid, sc, s_bx[wv],
                                                      my_scanned, my_probes, wsearch ? wmin : 0u));
                    StatAdd(ws->batched, 1);
                }
                if (lane == 0) lvl[d].off = off + kWave;
                WaveSync();
                continue;
            }
            if (d == last) {
                count = SatAdd(count, (unsigned long long)__popcll(m));
                if (kTuples && ok) {
                    unsigned long long slot = atomicAdd(&cnt->tuples, 1ull);
                    if (slot < tup_cap) {
                        uint32_t* o = tuples + slot * (uint64_t)(qv + 1);
                        for (uint32_t j = 0; j < last; j++) o[j] = matched[j];
                        o[last] = c;
                        o[qv] = ei;
                    } else {
                        atomicOr(&cnt->flags, kFlagTuples);
                    }
                }
                if (lane == 0) lvl[d].off = off + kWave;
                WaveSync();
                continue;
            }
            if (ok) surv[d * kWave + RankBelow(m)] = c;
            if (lane == 0) {
                lvl[d].off = off + kWave;
                lvl[d].cnt = (uint32_t)__popcll(m);
                lvl[d].pos = 0;
                lvl[d].end = limit;
            }
            WaveSync();
        }
        if (lane == 0) {
            if (count) pe[ei] = SatAdd(pe[ei], count);
            // Rows retire from `pending` a whole grab at a time, stolen items one by one.
            if (!from_row) atomicSub(&cnt->q_pending.v, 1ull);
            else if (r_next >= r_end) atomicSub(&cnt->q_pending.v, (unsigned long long)(r_end - r_first));
        }
    }
    {
        const long long now = (long long)Uniform64((unsigned long long)wall_clock64());
        if (idle) idle_sum += now - t_idle0;
        if (lane == 0) {
            atomicAdd(&cnt->wave_ticks, (unsigned long long)(now - t_born));
            atomicAdd(&cnt->idle_ticks, (unsigned long long)idle_sum);
        }
    }
    fold();
    WaveSync();
    if (lane == 0) {
        atomicAdd(&cnt->scanned, (unsigned long long)ws->scanned);
        atomicAdd(&cnt->probes, (unsigned long long)ws->probes);
        atomicAdd(&cnt->steps, (unsigned long long)ws->steps);
        atomicAdd(&cnt->step_lanes, (unsigned long long)ws->step_lanes);
        if (ws->batched) atomicAdd(&cnt->batched, (unsigned long long)ws->batched);
        if (ws->donated) atomicAdd(&cnt->donated, (unsigned long long)ws->donated);
    }
    flush_pe();
}

// Before a DFS launch: tickets taken but never filled in the previous launch
// leave head above tail; both restart at the larger value, so tickets (and
// the sequence numbers that publish them) are never reused.
__global__ void KQueueSync(Counters* __restrict__ cnt) {
    if (threadIdx.x == 0) {
        const unsigned long long h = cnt->q_head.v, t = cnt->q_tail.v, m = h > t ? h : t;
        cnt->q_head.v = m;
        cnt->q_tail.v = m;
    }
}

unsigned Grid1(uint64_t n, unsigned per, unsigned cap) {
    uint64_t g = (n + per - 1) / per;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

void LaunchSeed(const Tables& T, const EiList& L, uint32_t* rows, uint64_t cap_rows, Counters* cnt, hipStream_t s) {
    if (!L.n || !T.W) return;
    KSeed<<<Grid1(T.W, kWpb, 8192), kThreads, 0, s>>>(T, L, rows, cap_rows, cnt);
}

void LaunchGroupByTag(const uint32_t* in, uint64_t n, uint32_t words, uint32_t tag_shift, unsigned long long* hist,
                      uint32_t* out, hipStream_t s) {
    if (!n) return;
    const uint64_t per_block = 4096;
    const unsigned g = (unsigned)((n + per_block - 1) / per_block);
    if (words == 1) {
        KTagHist<1><<<g, kThreads, 0, s>>>(in, n, tag_shift, per_block, hist);
        KTagScan<<<1, kWave, 0, s>>>(hist);
        KTagScatter<1><<<g, kThreads, 0, s>>>(in, n, tag_shift, per_block, hist, out);
    } else {
        KTagHist<2><<<g, kThreads, 0, s>>>(in, n, tag_shift, per_block, hist);
        KTagScan<<<1, kWave, 0, s>>>(hist);
        KTagScatter<2><<<g, kThreads, 0, s>>>(in, n, tag_shift, per_block, hist, out);
    }
}

void LaunchSeed1(const Tables& T, const EiList& L, uint32_t* rows, uint64_t cap_rows, Counters* cnt, hipStream_t s) {
    if (!L.n || !T.W) return;
    KSeed1<<<Grid1(T.W, kWpb, 8192), kThreads, 0, s>>>(T, L, rows, cap_rows, cnt);
}

void LaunchExtend(const Tables& T, const uint32_t* in, uint64_t n_in, uint32_t depth, uint32_t* out, uint64_t cap_rows,
                  Counters* cnt, hipStream_t s) {
    if (!n_in) return;
    KExtend<<<Grid1(n_in, kWpb, 65535), kThreads, 0, s>>>(T, in, n_in, depth, out, cap_rows, cnt);
}

// LDS per final-DFS block (static tables + survivor stacks + cache) at nb
// blocks per CU: the CU's 160 KB less 4 KB, split evenly, in 1 KB steps
// (53 248 B at 3 blocks per CU, the largest block measured to keep full
// occupancy there).
constexpr size_t DfsBlockLds(int nb) { return ((size_t)(160 * 1024 - 4096) / (size_t)nb) & ~(size_t)1023; }

// LDS plan of a final-DFS launch: the register budget fixes the resident
// blocks per CU; the LDS those blocks leave after the static tables is split
// among their waves into survivor stacks and the stable-intersection cache.
template <bool kTuples>
DfsShape DfsLdsShapeT(uint32_t surv_levels) {
    static DfsShape cache[kMaxQV + 1] = {};
    static std::mutex mu;  // queries launch from several host threads
    std::lock_guard<std::mutex> lock(mu);
    if (surv_levels > (uint32_t)kMaxQV) surv_levels = kMaxQV;
    DfsShape& c = cache[surv_levels];
    if (c.blocks_per_cu) return c;
    int nb = 0;
    const size_t surv_bytes = (size_t)kWpb * surv_levels * kWave * sizeof(uint32_t);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, KDfs<kTuples>, kThreads, surv_bytes) != hipSuccess ||
        nb <= 0)
        throw HipError(hipErrorInvalidValue, "final-DFS occupancy query", __FILE__, __LINE__);
    // The cache takes what is left of DfsBlockLds per block.  The occupancy
    // query accepted 54.2 KB blocks at 3 per CU, but those ran 25 % slower
    // than 53.2 KB ones (measured), as if only 2 blocks fit: the budget is
    // the largest block measured to run at full occupancy.
    hipFuncAttributes attr{};
    if (hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(KDfs<kTuples>)) != hipSuccess)
        throw HipError(hipErrorInvalidValue, "final-DFS attributes", __FILE__, __LINE__);
    const size_t fixed = attr.sharedSizeBytes + surv_bytes;
    const size_t budget = DfsBlockLds(nb);
    uint32_t cap = budget > fixed ? (uint32_t)((budget - fixed) / (kWpb * sizeof(uint32_t))) : 0u;
    cap = cap > 4096 ? 4096 : cap & ~63u;
    for (;; cap = cap >= 64 ? cap - 64 : 0) {
        const size_t dyn = surv_bytes + (size_t)kWpb * cap * sizeof(uint32_t);
        int got = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&got, KDfs<kTuples>, kThreads, dyn) == hipSuccess &&
            got >= nb) {
            c.dyn_bytes = (uint32_t)dyn;
            c.stab_cap = cap;
            c.blocks_per_cu = (uint32_t)nb;
            return c;
        }
        if (!cap) throw HipError(hipErrorInvalidValue, "final-DFS LDS plan", __FILE__, __LINE__);
    }
}

DfsShape DfsLdsShape(uint32_t surv_levels, bool tuples) {
    return tuples ? DfsLdsShapeT<true>(surv_levels) : DfsLdsShapeT<false>(surv_levels);
}

uint32_t DfsBlocksPerCU() { return DfsLdsShape(kMaxQV, false).blocks_per_cu; }

void LaunchQueueSync(Counters* cnt, hipStream_t s) { KQueueSync<<<1, 64, 0, s>>>(cnt); }

void LaunchDfs(const DfsJob* jobs, uint32_t njobs, const unsigned long long* row_base, uint64_t n_rows, Counters* cnt,
               const StealQueue& Q, uint32_t* tuples, uint64_t tup_cap, uint32_t num_cus, uint32_t blocks_per_cu,
               uint32_t grab, uint32_t surv_levels, uint32_t cache_cap, hipStream_t s) {
    if (!n_rows || !njobs) return;
    DfsShape sh = DfsLdsShape(surv_levels, tuples != nullptr);
    if (cache_cap && cache_cap < sh.stab_cap) sh.stab_cap = cache_cap & ~63u;
    const uint32_t bpc = blocks_per_cu && blocks_per_cu < sh.blocks_per_cu ? blocks_per_cu : sh.blocks_per_cu;
    const unsigned g = Grid1(n_rows, kWpb, num_cus * bpc);
    grab = grab < 1 ? 1 : grab;
    if (tuples)
        KDfs<true><<<g, kThreads, sh.dyn_bytes, s>>>(jobs, njobs, row_base, grab, cnt, Q, tuples, tup_cap,
                                                      surv_levels, sh.stab_cap);
    else
        KDfs<false><<<g, kThreads, sh.dyn_bytes, s>>>(jobs, njobs, row_base, grab, cnt, Q, nullptr, 0, surv_levels,
                                                       sh.stab_cap);
}

}  // namespace cgd
