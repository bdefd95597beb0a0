// Device Stage 1 on gfx950: the CaLiG candidate index (lit bits) kept in HBM
// over a data graph shared by every query of a context, and the per-batch
// deltas it hands to Stage 2, produced on the device.  Replaces, per query,
//   CaLiG::constructCand / staticFilter        (index/calig.cpp:325-356, 705-717)
//   CaLiG::updateIndex + turnOnCond/turnOnCheck/addAndCheck/addAndCheckDelete/
//     delAndCheck/checkNei/tryNei               (index/calig.cpp:950-1069, 358-475, 536-703)
//   CaLiG::ConstructUpdate, mergeDeltas          (index/calig.cpp:776-804, 1071-1088)
//   CaLiGHelperGamma::UpdateVertexMappingCPU, PrepareUpdateAll/Index and the
//     host half of ConvertGlobalIndex            (index/calig_helper_gamma.cu:77-105, 238-445)
//
// State.  The graph (one per context): labels and sorted adjacency lists with
// slack ({off, len, cap} per vertex into a u32 pool), grown by the same
// per-list merge as the main index.  Per query and data vertex v one u32
//   vstate[v] = lit bits (0..11) | label code (12..15) | activated-this-batch (16..27)
// (label code = index of v's label among the query's labels, 15 = none), and
// per label-compatible pair (v, ui) a 16-bit "dead mask" over query vertices.
//
// Semantics.  Pair (v, ui) is lit iff it belongs to the greatest fixpoint of
//   check(v, ui, L): every query neighbour uj of ui has a data neighbour vj of v
//   with label(uj) and (vj, uj) in L, and for every group of ui's neighbours
//   sharing a label the candidate sets have a system of distinct
//   representatives (checkNei + tryNei, index/calig.cpp:358-408).
// For a lit pair, cand[ui][uj] is exactly {vj in N(v): (vj, uj) lit}, which is
// what the reference's sets hold for lit pairs after every batch; the index
// keys are built from that.  Lit sets only grow under insertions, so a batch
//   1. optimistically lights pairs that may have become supported (activation
//      closure, driven by dead masks), then
//   2. removes the optimistic pairs that fail the exact check (deletion rounds),
// and the result is the greatest fixpoint on the new graph:
//   * a pair's dead mask holds bit uj only while no lit neighbour supports uj
//     (bits are set when the pair dies, from the pairs alive in that round,
//     and cleared when a neighbour lights up or an edge arrives);
//   * a dead pair is re-lit (optimistically) when an event touches it and its
//     mask is empty.
// If a pair of the new fixpoint stayed dark, take the one that died earliest:
// its mask is empty (every supporter was alive when it died, or lit up or got
// its edge later, clearing the bit) and no event touched it since (else it
// would have been re-lit), so it passed the exact check against the pairs
// alive when it died -- a contradiction.  The oracle's from-scratch fixpoint
// pins the result batch by batch.
//
// Work lists hold pairs as (v << 4) | ui.  Kernels that walk adjacency lists
// lay the lists of a whole pair list end to end over the lanes (an exclusive
// scan of their lengths, 64 list entries per wavefront step), so a hub does
// not stall the other 63 lanes.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "api.h"
#include "wave.h"

namespace cgd {

constexpr uint32_t kCodeNone = 15;
constexpr uint32_t kSmallDeg = 48;  // pairs with longer lists are checked a wavefront each

// Per-query constant tables, passed by value.
struct QTab {
    uint32_t qv, ndl;
    int32_t dlabel[kMaxQV];        // distinct query labels; code = index
    uint8_t qcode[kMaxQV];         // label code of each query vertex
    uint16_t nbr[kMaxQV];          // query neighbours of ui (bit uj)
    uint16_t by_code[16];          // query vertices with label code c
    uint16_t uj_of[kMaxQV][16];    // neighbours of ui with label code c
    uint16_t grp[kMaxQV];          // neighbours of ui that share their label with another neighbour of ui
    uint64_t gneed[kMaxQV];        // 4-bit per uj: size of uj's label group among ui's neighbours
    uint8_t eidx[kMaxQV][kMaxQV];  // directed query edge or 0xFF
    uint32_t ushift;               // key format: dir << ushift | owner << 32 | target
};

struct GView {
    const MRec* rec;
    const uint32_t* pool;
    uint32_t n;
};

struct SView {
    uint32_t* vstate;
    uint32_t* queued;  // bit ui: (v, ui) is on the list being built
    uint32_t* dead;    // 16-bit dead masks, two per word, index v * qv + ui
    uint32_t* new_of;  // compressed id or ~0u
    uint32_t* old_of;
};

struct alignas(128) S1Cnt {
    unsigned long long n_list[6];   // work-list sizes, one per work list L[0..5] (see S1Query)
    unsigned long long n_died, n_big, n_keys, n_upd, ovf, items, scan_total, newv;
    unsigned long long dirmask[2];  // directed query edges that received an update key
    unsigned long long pairs;       // label-compatible pairs (attach)
    unsigned long long n_killed;    // pairs lit before the batch that died in it
};

namespace {

constexpr int kThreads = 256;
constexpr int kWpb = kThreads / kWave;

__device__ __forceinline__ uint32_t Lit(uint32_t s) { return s & 0xFFFu; }
__device__ __forceinline__ uint32_t Code(uint32_t s) { return (s >> 12) & 15u; }
__device__ __forceinline__ uint32_t Opt(uint32_t s) { return (s >> 16) & 0xFFFu; }

__device__ __forceinline__ uint32_t DeadGet(const uint32_t* dead, uint32_t v, uint32_t ui, uint32_t qv) {
    const uint64_t i = (uint64_t)v * qv + ui;
    return (dead[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu;
}
__device__ __forceinline__ void DeadClear(uint32_t* dead, uint32_t v, uint32_t ui, uint32_t qv, uint32_t bits) {
    const uint64_t i = (uint64_t)v * qv + ui;
    atomicAnd(&dead[i >> 1], ~(bits << ((i & 1) * 16)));
}
// Only one writer per pair per launch, and no concurrent DeadClear (other launch).
__device__ __forceinline__ void DeadSet(uint32_t* dead, uint32_t v, uint32_t ui, uint32_t qv, uint32_t bits) {
    const uint64_t i = (uint64_t)v * qv + ui;
    const uint32_t sh = (uint32_t)(i & 1) * 16;
    atomicAnd(&dead[i >> 1], ~(0xFFFFu << sh));
    if (bits) atomicOr(&dead[i >> 1], bits << sh);
}

// Per-lane slots on a shared counter for `cnt` items (all lanes call).
__device__ __forceinline__ uint64_t WaveReserve(uint32_t cnt, unsigned long long* counter) {
    const uint32_t incl = WaveInclusiveScan(cnt);
    const uint32_t total = ReadLane(incl, 63);
    unsigned long long base = 0;
    if (total && LaneId() == 63) base = atomicAdd(counter, (unsigned long long)total);
    base = ReadLane64(base, 63);
    return base + incl - cnt;
}

// Marks (v, u) queued; true for the lane that queued it first.
__device__ __forceinline__ bool Enqueue(uint32_t* queued, uint32_t v, uint32_t u) {
    return !((atomicOr(&queued[v], 1u << u) >> u) & 1u);
}

// ---------------------------------------------------------------- load-balanced list walks
// The items of a pair list are the entries of its pairs' adjacency lists,
// offs[p] = first item of pair p (exclusive scan of list lengths, offs[np] =
// total).  Locates item it0 + lane: its pair and position in that list.
struct LbItem {
    uint64_t p;
    uint32_t j;
    bool valid;
};

__device__ __forceinline__ LbItem LbLocate(const uint64_t* __restrict__ offs, uint64_t np, uint64_t it0,
                                           uint64_t total) {
    const uint32_t lane = LaneId();
    const uint64_t it = it0 + lane;
    // p0: the last pair starting at or before it0 (uniform).
    uint64_t lo = 0, hi = np;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (offs[mid] <= it0) lo = mid;
        else hi = mid;
    }
    const uint64_t p0 = lo;
    const uint64_t e = p0 + 1 + lane <= np ? offs[p0 + 1 + lane] : ~0ull;
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1) {
        const uint64_t ee = __shfl(e, (int)(pos + step - 1), kWave);
        if (ee <= it) pos += step;
    }
    const uint64_t last = ReadLane64(e, 63);
    const uint64_t prev = __shfl(e, (int)(pos ? pos - 1 : 0), kWave);
    LbItem r;
    r.valid = it < total;
    if (last <= it) {
        // Beyond the 64 pairs after p0 (runs of empty lists): search the rest.
        uint64_t a = p0 + 64, b = np;
        while (b - a > 1) {
            const uint64_t mid = (a + b) >> 1;
            if (offs[mid] <= it) a = mid;
            else b = mid;
        }
        r.p = a;
        r.j = (uint32_t)(it - offs[a]);
    } else {
        r.p = p0 + pos;
        r.j = (uint32_t)(it - (pos ? prev : offs[p0]));
    }
    if (!r.valid) {
        r.p = 0;
        r.j = 0;
    }
    return r;
}

// ---------------------------------------------------------------- graph kernels
// Both orientations of every valid pair of stream ints [start, start + 2 * np)
// as keys (src << vbits) | dst (invalid pairs: ~0, sorted to the end).
// Both directions of every pair of the batch: insertions into keys, and
// deletions -- a pair of negative ids (-x-1, -y-1) removes edge (x, y)
// (index/calig.cpp:980-989) -- into dkeys; ~0 for the other kind and for
// pairs that name no valid edge.
__global__ void KGraphPairKeys(const int32_t* __restrict__ ints, uint64_t start, uint64_t np, uint32_t n,
                               uint32_t vbits, uint64_t* __restrict__ keys, uint64_t* __restrict__ dkeys) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t a = ints[start + 2 * i], b = ints[start + 2 * i + 1];
        const bool ok = a >= 0 && b >= 0 && (uint32_t)a < n && (uint32_t)b < n && a != b;
        keys[2 * i] = ok ? ((uint64_t)a << vbits) | (uint32_t)b : ~0ull;
        keys[2 * i + 1] = ok ? ((uint64_t)b << vbits) | (uint32_t)a : ~0ull;
        const int64_t x = -(int64_t)a - 1, y = -(int64_t)b - 1;
        const bool del = a < 0 && b < 0 && x < (int64_t)n && y < (int64_t)n && x != y;
        dkeys[2 * i] = del ? ((uint64_t)x << vbits) | (uint64_t)y : ~0ull;
        dkeys[2 * i + 1] = del ? ((uint64_t)y << vbits) | (uint64_t)x : ~0ull;
    }
}

// keep[i] = sorted unique key i is a valid edge absent from (present: in)
// the graph.
__global__ void KGraphAbsent(const uint64_t* __restrict__ keys, const unsigned long long* __restrict__ n_dev,
                             uint64_t n_upper, uint32_t vbits, GView G, uint32_t present, uint32_t* __restrict__ keep) {
    const uint64_t n = *n_dev;
    const uint64_t vmask = (1ull << vbits) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_upper;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t k = 0;
        if (i < n && keys[i] != ~0ull) {
            const uint32_t a = (uint32_t)(keys[i] >> vbits), b = (uint32_t)(keys[i] & vmask);
            const MRec r = G.rec[a];
            k = SortedContains(G.pool + r.off, r.len, b) == (present != 0) ? 1u : 0u;
        }
        keep[i] = k;
    }
}

__global__ void KCompactKeys(const uint64_t* __restrict__ keys, uint64_t n_upper, const uint32_t* __restrict__ keep,
                             const uint32_t* __restrict__ pos, uint64_t* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_upper;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (keep[i]) out[pos[i]] = keys[i];
}

__global__ void KEdgeKeys(const uint32_t* __restrict__ edges, uint64_t m, uint32_t n, uint32_t vbits,
                          uint64_t* __restrict__ keys) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t a = edges[2 * i], b = edges[2 * i + 1];
        const bool ok = a < n && b < n && a != b;
        keys[2 * i] = ok ? ((uint64_t)a << vbits) | b : ~0ull;
        keys[2 * i + 1] = ok ? ((uint64_t)b << vbits) | a : ~0ull;
    }
}

// ---------------------------------------------------------------- per-query kernels
// vstate = label code (and, for the initial build, every compatible pair lit).
__global__ void KS1Codes(const int32_t* __restrict__ labels, uint32_t n, QTab Q, uint32_t* __restrict__ vstate,
                         uint32_t light_all, unsigned long long* __restrict__ pairs) {
    unsigned long long mine = 0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t L = labels[v];
        uint32_t c = kCodeNone;
        for (uint32_t k = 0; k < Q.ndl; k++)
            if (Q.dlabel[k] == L) c = k;
        const uint32_t bc = c == kCodeNone ? 0u : Q.by_code[c];
        vstate[v] = (c << 12) | (light_all ? bc : 0u);
        mine += (unsigned long long)__popc(bc);
    }
    const unsigned long long tot = ReadLane64(WaveInclusiveScan(mine), 63);
    if (LaneId() == 0 && tot) atomicAdd(pairs, tot);
}

// Every pair (v, ui) with bit ui in (lit_only ? lit : label-compatible) bits.
__global__ void KS1ListPairs(const uint32_t* __restrict__ vstate, uint32_t n, QTab Q, uint32_t lit_only,
                             uint32_t* __restrict__ out, unsigned long long* __restrict__ n_out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint64_t v = base + threadIdx.x;
        uint32_t bits = 0;
        if (v < n) {
            const uint32_t s = vstate[v];
            bits = lit_only ? Lit(s) : (Code(s) == kCodeNone ? 0u : Q.by_code[Code(s)]);
        }
        const uint64_t at = WaveReserve((uint32_t)__popc(bits), n_out);
        uint32_t k = 0;
        while (bits) {
            const uint32_t u = __ffs(bits) - 1;
            bits &= bits - 1;
            out[at + k++] = ((uint32_t)v << 4) | u;
        }
    }
}

// Hall's condition for the members of `S` (bit = query vertex) whose
// candidate elements the wavefront gathered in lists[uj][0..len[uj]).
// Returns true when a system of distinct representatives exists.
__device__ bool HallHolds(uint32_t S, const uint32_t (*lists)[kMaxQV], const uint32_t* len, uint32_t* elems,
                          uint64_t (*masks)[2]) {
    const uint32_t lane = LaneId();
    uint32_t mem[kMaxQV];
    uint32_t nm = 0;
    for (uint32_t u = 0; u < (uint32_t)kMaxQV; u++)
        if ((S >> u) & 1u) mem[nm++] = u;
    if (lane == 0) {
        uint32_t ne = 0;
        for (uint32_t a = 0; a < nm; a++) {
            masks[a][0] = masks[a][1] = 0;
            const uint32_t u = mem[a];
            for (uint32_t k = 0; k < len[u]; k++) {
                const uint32_t x = lists[u][k];
                uint32_t id = ne;
                for (uint32_t t = 0; t < ne; t++)
                    if (elems[t] == x) {
                        id = t;
                        break;
                    }
                if (id == ne) elems[ne++] = x;
                masks[a][id >> 6] |= 1ull << (id & 63);
            }
        }
    }
    WaveSync();
    bool ok = true;
    const uint32_t nsub = 1u << nm;
    for (uint32_t x = lane + 1; x < nsub; x += kWave) {
        uint64_t m0 = 0, m1 = 0;
        for (uint32_t a = 0; a < nm; a++)
            if ((x >> a) & 1u) {
                m0 |= masks[a][0];
                m1 |= masks[a][1];
            }
        if ((uint32_t)(__popcll(m0) + __popcll(m1)) < (uint32_t)__popc(x)) ok = false;
    }
    return __ballot(!ok) == 0;
}

// Exact check of list pairs, one lane per pair with a short adjacency list.
// Failing pairs go to `died` with their dead mask (query neighbours without
// a lit supporter); pairs with long lists, or whose distinct-representative
// condition the counts do not settle, go to `big`.
__global__ void __launch_bounds__(kThreads) KS1Check(GView G, QTab Q, SView S, const uint32_t* __restrict__ W,
                                                     const unsigned long long* __restrict__ n_dev,
                                                     uint32_t* __restrict__ died, uint16_t* __restrict__ died_bits,
                                                     unsigned long long* __restrict__ n_died,
                                                     uint32_t* __restrict__ big,
                                                     unsigned long long* __restrict__ n_big) {
    const uint64_t n = *n_dev;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool fail = false, tobig = false;
        uint32_t p = 0, dbits = 0;
        if (i < n) {
            p = W[i];
            const uint32_t v = p >> 4, ui = p & 15u;
            atomicAnd(&S.queued[v], ~(1u << ui));
            const MRec r = G.rec[v];
            if (r.len > kSmallDeg) {
                tobig = true;
            } else {
                const uint32_t* A = G.pool + r.off;
                const uint32_t need = Q.nbr[ui], grp = Q.grp[ui];
                uint32_t has = 0;
                uint64_t cnt = 0;  // 4-bit saturating counts of group members
                for (uint32_t k = 0; k < r.len; k++) {
                    const uint32_t s = S.vstate[A[k]];
                    const uint32_t c = Code(s);
                    const uint32_t sup = c == kCodeNone ? 0u : (Lit(s) & Q.uj_of[ui][c]);
                    has |= sup;
                    uint32_t g = sup & grp;
                    while (g) {
                        const uint32_t u = __ffs(g) - 1;
                        g &= g - 1;
                        if (((cnt >> (4 * u)) & 15u) != 15u) cnt += 1ull << (4 * u);
                    }
                }
                if ((has & need) != need) {
                    fail = true;
                    dbits = need & ~has;
                } else if (grp) {
                    uint32_t g = grp;
                    while (g) {
                        const uint32_t u = __ffs(g) - 1;
                        g &= g - 1;
                        if (((cnt >> (4 * u)) & 15u) < ((Q.gneed[ui] >> (4 * u)) & 15u)) tobig = true;
                    }
                }
            }
        }
        const uint64_t at = WaveReserve(fail ? 1u : 0u, n_died);
        if (fail) {
            died[at] = p;
            died_bits[at] = (uint16_t)dbits;
        }
        const uint64_t bt = WaveReserve(tobig ? 1u : 0u, n_big);
        if (tobig) big[bt] = p;
    }
}

// Exact check, one wavefront per pair (long lists and undecided groups).
__global__ void __launch_bounds__(kThreads) KS1CheckWave(GView G, QTab Q, SView S, const uint32_t* __restrict__ B,
                                                         const unsigned long long* __restrict__ n_dev,
                                                         uint32_t* __restrict__ died, uint16_t* __restrict__ died_bits,
                                                         unsigned long long* __restrict__ n_died) {
    __shared__ uint32_t lists[kWpb][kMaxQV][kMaxQV];
    __shared__ uint32_t lens[kWpb][kMaxQV];
    __shared__ uint32_t elems[kWpb][kMaxQV * kMaxQV];
    __shared__ uint64_t masks[kWpb][kMaxQV][2];
    const uint32_t lane = LaneId(), w = threadIdx.x / kWave;
    const uint64_t n = *n_dev;
    const uint64_t nw = (uint64_t)gridDim.x * kWpb;
    for (uint64_t i = (uint64_t)blockIdx.x * kWpb + w; i < n; i += nw) {
        const uint32_t p = B[i];
        const uint32_t v = p >> 4, ui = p & 15u;
        const MRec r = G.rec[v];
        const uint32_t* A = G.pool + r.off;
        const uint32_t need = Q.nbr[ui], grp = Q.grp[ui];
        uint32_t has = 0;
        uint64_t cnt = 0;
        for (uint32_t k0 = 0; k0 < r.len; k0 += kWave) {
            const uint32_t k = k0 + lane;
            uint32_t sup = 0;
            if (k < r.len) {
                const uint32_t s = S.vstate[A[k]];
                const uint32_t c = Code(s);
                sup = c == kCodeNone ? 0u : (Lit(s) & Q.uj_of[ui][c]);
            }
            uint32_t nb = need;
            while (nb) {
                const uint32_t u = __ffs(nb) - 1;
                nb &= nb - 1;
                const uint64_t b = __ballot((sup >> u) & 1u);
                if (b) {
                    has |= 1u << u;
                    if ((grp >> u) & 1u) {
                        const uint64_t c = ((cnt >> (4 * u)) & 15u) + (uint64_t)__popcll(b);
                        cnt = (cnt & ~(15ull << (4 * u))) | ((c > 15 ? 15ull : c) << (4 * u));
                    }
                }
            }
        }
        bool fail = false;
        uint32_t dbits = 0;
        if ((has & need) != need) {
            fail = true;
            dbits = need & ~has;
        } else if (grp) {
            uint32_t small = 0, g = grp;
            while (g) {
                const uint32_t u = __ffs(g) - 1;
                g &= g - 1;
                if (((cnt >> (4 * u)) & 15u) < ((Q.gneed[ui] >> (4 * u)) & 15u)) small |= 1u << u;
            }
            if (small) {
                // Gather the small members' candidates (each fewer than its
                // group size) and test Hall's condition on them; members
                // with at least as many candidates as their group always
                // find a free one.
                if (lane < (uint32_t)kMaxQV) lens[w][lane] = 0;
                WaveSync();
                for (uint32_t k0 = 0; k0 < r.len; k0 += kWave) {
                    const uint32_t k = k0 + lane;
                    uint32_t sup = 0, x = 0;
                    if (k < r.len) {
                        x = A[k];
                        const uint32_t s = S.vstate[x];
                        const uint32_t c = Code(s);
                        sup = c == kCodeNone ? 0u : (Lit(s) & Q.uj_of[ui][c] & small);
                    }
                    uint32_t sm = small;
                    while (sm) {
                        const uint32_t u = __ffs(sm) - 1;
                        sm &= sm - 1;
                        const uint64_t b = __ballot((sup >> u) & 1u);
                        if ((sup >> u) & 1u) {
                            const uint32_t at = lens[w][u] + RankBelow(b);
                            if (at < (uint32_t)kMaxQV) lists[w][u][at] = x;
                        }
                        WaveSync();
                        if (lane == 0) lens[w][u] += (uint32_t)__popcll(b);
                        WaveSync();
                    }
                }
                fail = !HallHolds(small, lists[w], lens[w], elems[w], masks[w]);
                WaveSync();
            }
        }
        if (fail && lane == 0) {
            const unsigned long long at = atomicAdd(n_died, 1ull);
            died[at] = p;
            died_bits[at] = (uint16_t)dbits;
        }
    }
}

// Deaths of a round: clear the lit (and activated) bit, record the dead mask.
// With `killed`, deaths of pairs lit before the batch (possible only when
// the batch deletes edges) are also listed there and marked in bits 16..27
// of `queued` (their index entries leave the main index).
__global__ void KS1Apply(QTab Q, SView S, const uint32_t* __restrict__ died, const uint16_t* __restrict__ bits,
                         const unsigned long long* __restrict__ n_dev, uint32_t* __restrict__ killed,
                         unsigned long long* __restrict__ n_killed) {
    const uint64_t n = *n_dev;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool old_pair = false;
        uint32_t p = 0;
        if (i < n) {
            p = died[i];
            const uint32_t v = p >> 4, ui = p & 15u;
            const uint32_t old = atomicAnd(&S.vstate[v], ~((1u << ui) | (1u << (16 + ui))));
            DeadSet(S.dead, v, ui, Q.qv, bits[i]);
            if (killed && !((Opt(old) >> ui) & 1u)) {
                old_pair = true;
                atomicOr(&S.queued[v], 1u << (16 + ui));
            }
        }
        if (killed) {
            const uint64_t at = WaveReserve(old_pair ? 1u : 0u, n_killed);
            if (old_pair) killed[at] = p;
        }
    }
}

// Deleted edge a -> b: lit pairs (a, ui) from before the batch with a query
// neighbour labelled like b may have lost their support; queued for the
// deletion rounds (pairs lit in this batch are on the list already).
__global__ void KS1SeedDeleted(QTab Q, SView S, const uint64_t* __restrict__ dkeys, uint64_t nd, uint32_t vbits,
                               uint32_t* __restrict__ out, unsigned long long* __restrict__ n_out) {
    const uint64_t vmask = (1ull << vbits) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nd; base += stride) {
        const uint64_t i = base + threadIdx.x;
        uint32_t a = 0, cand = 0;
        if (i < nd) {
            a = (uint32_t)(dkeys[i] >> vbits);
            const uint32_t sa = S.vstate[a], cb = Code(S.vstate[(uint32_t)(dkeys[i] & vmask)]);
            if (cb != kCodeNone) {
                uint32_t l = Lit(sa) & ~Opt(sa);
                while (l) {
                    const uint32_t u = __ffs(l) - 1;
                    l &= l - 1;
                    if (Q.uj_of[u][cb]) cand |= 1u << u;
                }
            }
        }
        for (uint32_t u = 0; u < Q.qv; u++) {
            const bool want = ((cand >> u) & 1u) && Enqueue(S.queued, a, u);
            const uint64_t at = WaveReserve(want ? 1u : 0u, n_out);
            if (want) out[at] = (a << 4) | u;
        }
    }
}

// Deleted edge a -> b: the main-index entries it carried, a -> b of
// direction (ui, uj) for every pair (a, ui), (b, uj) lit before the batch
// (the reverse entries come from the key b -> a).  Run before the deletion
// rounds, while lit & ~activated is still the lit set before the batch.
__global__ void KS1DelEdgeKeys(QTab Q, SView S, const uint64_t* __restrict__ dkeys, uint64_t nd, uint32_t vbits,
                               uint64_t* __restrict__ keys, uint64_t cap, unsigned long long* __restrict__ n_keys,
                               unsigned long long* __restrict__ ovf, unsigned long long* __restrict__ dirmask) {
    const uint64_t vmask = (1ull << vbits) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long dm0 = 0, dm1 = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nd; base += stride) {
        const uint64_t i = base + threadIdx.x;
        uint32_t a = 0, b = 0, la = 0, lb = 0, cb = kCodeNone, cnt = 0;
        if (i < nd) {
            a = (uint32_t)(dkeys[i] >> vbits);
            b = (uint32_t)(dkeys[i] & vmask);
            const uint32_t sa = S.vstate[a], sb = S.vstate[b];
            cb = Code(sb);
            la = Lit(sa) & ~Opt(sa);
            lb = Lit(sb) & ~Opt(sb);
            if (cb != kCodeNone) {
                uint32_t l = la;
                while (l) {
                    const uint32_t u = __ffs(l) - 1;
                    l &= l - 1;
                    cnt += (uint32_t)__popc(Q.uj_of[u][cb] & lb);
                }
            }
        }
        uint64_t at = WaveReserve(cnt, n_keys);
        if (cnt) {
            if (at + cnt > cap) {
                atomicOr(ovf, 1ull);
            } else {
                const uint64_t na = S.new_of[a], nb = S.new_of[b];
                uint32_t l = la;
                while (l) {
                    const uint32_t ui = __ffs(l) - 1;
                    l &= l - 1;
                    uint32_t m = Q.uj_of[ui][cb] & lb;
                    while (m) {
                        const uint32_t uj = __ffs(m) - 1;
                        m &= m - 1;
                        const uint32_t d = Q.eidx[ui][uj];
                        keys[at++] = ((uint64_t)d << Q.ushift) | (na << 32) | nb;
                        if (d < 64) dm0 |= 1ull << d;
                        else dm1 |= 1ull << (d - 64);
                    }
                }
            }
        }
    }
    if (dm0) atomicOr(&dirmask[0], dm0);
    if (dm1) atomicOr(&dirmask[1], dm1);
}

// Pairs lit before the batch that died in it: the main-index entries they
// shared with the pairs lit before the batch, both directions, over the
// current adjacency (entries over deleted edges come from KS1DelEdgeKeys).
// When both ends died, the end with the smaller vertex emits them.
__global__ void __launch_bounds__(kThreads) KS1EmitKilled(GView G, QTab Q, SView S, const uint32_t* __restrict__ K,
                                                          const uint64_t* __restrict__ offs, uint64_t np,
                                                          uint64_t* __restrict__ keys, uint64_t cap,
                                                          unsigned long long* __restrict__ n_keys,
                                                          unsigned long long* __restrict__ ovf) {
    const uint64_t total = offs[np];
    const uint64_t nw = (uint64_t)gridDim.x * kWpb;
    for (uint64_t t = (uint64_t)blockIdx.x * kWpb + threadIdx.x / kWave; t * kWave < total; t += nw) {
        const LbItem it = LbLocate(offs, np, t * kWave, total);
        uint32_t m = 0, v = 0, ui = 0, vj = 0;
        if (it.valid) {
            const uint32_t p = K[it.p];
            v = p >> 4;
            ui = p & 15u;
            vj = G.pool[G.rec[v].off + it.j];
            const uint32_t s = S.vstate[vj];
            const uint32_t c = Code(s);
            const uint32_t kb = (S.queued[vj] >> 16) & 0xFFFu;
            m = c == kCodeNone ? 0u : (Q.uj_of[ui][c] & ((Lit(s) & ~Opt(s)) | kb));
            if (vj < v) m &= ~kb;
        }
        const uint32_t cnt = 2u * (uint32_t)__popc(m);
        uint64_t at = WaveReserve(cnt, n_keys);
        if (cnt) {
            if (at + cnt > cap) {
                atomicOr(ovf, 1ull);
            } else {
                const uint64_t nu = S.new_of[v], nv = S.new_of[vj];
                while (m) {
                    const uint32_t uj = __ffs(m) - 1;
                    m &= m - 1;
                    keys[at++] = ((uint64_t)Q.eidx[ui][uj] << Q.ushift) | (nu << 32) | nv;
                    keys[at++] = ((uint64_t)Q.eidx[uj][ui] << Q.ushift) | (nv << 32) | nu;
                }
            }
        }
    }
}

__global__ void KS1ClearKilled(SView S, const uint32_t* __restrict__ K, const unsigned long long* __restrict__ n_dev) {
    const uint64_t n = *n_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = K[i];
        atomicAnd(&S.queued[p >> 4], ~(1u << (16 + (p & 15u))));
    }
}

// offs[i] = adjacency length of the vertex of pair i (before the scan); with
// n_dev, entries from *n_dev up to n_upper are 0 (pairs the list does not hold).
__global__ void KS1Deg(GView G, const uint32_t* __restrict__ L, uint64_t n_upper,
                       const unsigned long long* __restrict__ n_dev, uint64_t* __restrict__ offs) {
    const uint64_t n = n_dev ? *n_dev : n_upper;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_upper;
         i += (uint64_t)gridDim.x * blockDim.x)
        offs[i] = i < n ? G.rec[L[i] >> 4].len : 0u;
}

// After deaths: queue the lit neighbour pairs they supported (only the pairs
// lit in this batch when `opt_only`: pairs lit before the batch stay in the
// new fixpoint).
__global__ void __launch_bounds__(kThreads) KS1Expand(GView G, QTab Q, SView S, const uint32_t* __restrict__ D,
                                                      const uint64_t* __restrict__ offs, uint64_t np,
                                                      uint32_t opt_only, uint32_t* __restrict__ out,
                                                      unsigned long long* __restrict__ n_out) {
    const uint64_t total = offs[np];
    const uint64_t nw = (uint64_t)gridDim.x * kWpb;
    for (uint64_t t = (uint64_t)blockIdx.x * kWpb + threadIdx.x / kWave; t * kWave < total; t += nw) {
        const LbItem it = LbLocate(offs, np, t * kWave, total);
        uint32_t cand = 0, vj = 0;
        if (it.valid) {
            const uint32_t p = D[it.p], v = p >> 4, ui = p & 15u;
            vj = G.pool[G.rec[v].off + it.j];
            const uint32_t s = S.vstate[vj];
            const uint32_t c = Code(s);
            cand = c == kCodeNone ? 0u : (Q.uj_of[ui][c] & Lit(s) & (opt_only ? Opt(s) : 0xFFFu));
        }
        for (uint32_t u = 0; u < Q.qv; u++) {
            const bool want = ((cand >> u) & 1u) && Enqueue(S.queued, vj, u);
            const uint64_t at = WaveReserve(want ? 1u : 0u, n_out);
            if (want) out[at] = (vj << 4) | u;
        }
    }
}

// New edge (a, b): every dark pair (a, ui) with a neighbour uj labelled like b
// may now be supported: clear those bits of its dead mask and touch it.
__global__ void KS1TouchNew(QTab Q, SView S, const uint64_t* __restrict__ ekeys, uint64_t ne, uint32_t vbits,
                            uint32_t* __restrict__ out, unsigned long long* __restrict__ n_out) {
    const uint64_t vmask = (1ull << vbits) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < ne; base += stride) {
        const uint64_t i = base + threadIdx.x;
        uint32_t a = 0, cand = 0, cb = kCodeNone;
        if (i < ne) {
            a = (uint32_t)(ekeys[i] >> vbits);
            const uint32_t b = (uint32_t)(ekeys[i] & vmask);
            const uint32_t sa = S.vstate[a];
            cb = Code(S.vstate[b]);
            if (Code(sa) != kCodeNone && cb != kCodeNone) cand = Q.by_code[Code(sa)] & ~Lit(sa);
        }
        for (uint32_t u = 0; u < Q.qv; u++) {
            bool want = false;
            if ((cand >> u) & 1u) {
                const uint32_t m = Q.uj_of[u][cb];
                if (m) {
                    DeadClear(S.dead, a, u, Q.qv, m);
                    want = Enqueue(S.queued, a, u);
                }
            }
            const uint64_t at = WaveReserve(want ? 1u : 0u, n_out);
            if (want) out[at] = (a << 4) | u;
        }
    }
}

// Touched pairs that are still dark with an empty dead mask light up
// (optimistically; the deletion rounds check them).
__global__ void KS1Activate(QTab Q, SView S, const uint32_t* __restrict__ T, const unsigned long long* __restrict__ n_dev,
                            uint32_t* __restrict__ act, unsigned long long* __restrict__ n_act,
                            uint32_t* __restrict__ opt, unsigned long long* __restrict__ n_opt) {
    const uint64_t n = *n_dev;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool on = false;
        uint32_t p = 0;
        if (i < n) {
            p = T[i];
            const uint32_t v = p >> 4, u = p & 15u;
            atomicAnd(&S.queued[v], ~(1u << u));
            if (!((Lit(S.vstate[v]) >> u) & 1u) && DeadGet(S.dead, v, u, Q.qv) == 0) {
                atomicOr(&S.vstate[v], (1u << u) | (1u << (16 + u)));
                on = true;
            }
        }
        const uint64_t at = WaveReserve(on ? 1u : 0u, n_act);
        if (on) act[at] = p;
        const uint64_t ot = WaveReserve(on ? 1u : 0u, n_opt);
        if (on) opt[ot] = p;
    }
}

// A pair lit up: its dark neighbour pairs lose the dead bit for it and are touched.
__global__ void __launch_bounds__(kThreads) KS1Push(GView G, QTab Q, SView S, const uint32_t* __restrict__ A,
                                                    const uint64_t* __restrict__ offs, uint64_t np,
                                                    uint32_t* __restrict__ out, unsigned long long* __restrict__ n_out) {
    const uint64_t total = offs[np];
    const uint64_t nw = (uint64_t)gridDim.x * kWpb;
    for (uint64_t t = (uint64_t)blockIdx.x * kWpb + threadIdx.x / kWave; t * kWave < total; t += nw) {
        const LbItem it = LbLocate(offs, np, t * kWave, total);
        uint32_t cand = 0, vj = 0, ui = 0;
        if (it.valid) {
            const uint32_t p = A[it.p], v = p >> 4;
            ui = p & 15u;
            vj = G.pool[G.rec[v].off + it.j];
            const uint32_t s = S.vstate[vj];
            const uint32_t c = Code(s);
            cand = c == kCodeNone ? 0u : (Q.uj_of[ui][c] & ~Lit(s));
        }
        for (uint32_t u = 0; u < Q.qv; u++) {
            bool want = false;
            if ((cand >> u) & 1u) {
                DeadClear(S.dead, vj, u, Q.qv, 1u << ui);
                want = Enqueue(S.queued, vj, u);
            }
            const uint64_t at = WaveReserve(want ? 1u : 0u, n_out);
            if (want) out[at] = (vj << 4) | u;
        }
    }
}

// Pairs of `opt` still lit (the batch's new pairs) -> Y; their vertices
// without a compressed id yet are new to the compressed space (a vertex
// keeps its id when deletions leave it without lit pairs).
__global__ void KS1Survivors(SView S, const uint32_t* __restrict__ opt, const unsigned long long* __restrict__ n_dev,
                             uint32_t* __restrict__ Y, unsigned long long* __restrict__ n_y,
                             uint32_t* __restrict__ newv) {
    const uint64_t n = *n_dev;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t p = 0;
        if (i < n) {
            p = opt[i];
            const uint32_t v = p >> 4, u = p & 15u;
            const uint32_t s = S.vstate[v];
            if ((Opt(s) >> u) & 1u) {
                keep = true;
                if (S.new_of[v] == ~0u) atomicOr(&newv[v >> 5], 1u << (v & 31));
            }
        }
        const uint64_t at = WaveReserve(keep ? 1u : 0u, n_y);
        if (keep) Y[at] = p;
    }
}

// Initial build: every vertex with a lit bit is new.
__global__ void KS1LitBitmap(const uint32_t* __restrict__ vstate, uint32_t n, uint32_t* __restrict__ newv) {
    const uint32_t nw = (n + 31) / 32;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t bits = 0;
        for (uint32_t k = 0; k < 32; k++) {
            const uint64_t v = w * 32 + k;
            if (v < n && Lit(vstate[v])) bits |= 1u << k;
        }
        newv[w] = bits;
    }
}

__global__ void KPopc(const uint32_t* __restrict__ bits, uint64_t nw, uint32_t* __restrict__ out) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * blockDim.x)
        out[w] = (uint32_t)__popc(bits[w]);
}

// New compressed ids in ascending original id (UpdateVertexMappingCPU's rule
// with a deterministic order); clears the bitmap for the next batch.
__global__ void KS1AssignIds(uint32_t* __restrict__ newv, const uint32_t* __restrict__ prefix, uint64_t nw,
                             uint32_t dv0, SView S) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t bits = newv[w];
        uint32_t id = dv0 + prefix[w];
        while (bits) {
            const uint32_t k = __ffs(bits) - 1;
            bits &= bits - 1;
            const uint32_t v = (uint32_t)(w * 32 + k);
            S.new_of[v] = id;
            S.old_of[id] = v;
            id++;
        }
        newv[w] = 0;
    }
}

// Index keys of pairs Y: for every lit neighbour pair (vj, uj) of (v, ui) the
// entry v -> vj of direction (ui, uj); with `both`, also vj -> v of (uj, ui)
// (a batch's new pairs: their neighbours' lists gain them).
__global__ void __launch_bounds__(kThreads) KS1EmitPairs(GView G, QTab Q, SView S, const uint32_t* __restrict__ Y,
                                                         const uint64_t* __restrict__ offs, uint64_t np, uint32_t both,
                                                         uint64_t* __restrict__ keys, uint64_t cap,
                                                         unsigned long long* __restrict__ n_keys,
                                                         unsigned long long* __restrict__ ovf) {
    const uint64_t total = offs[np];
    const uint64_t nw = (uint64_t)gridDim.x * kWpb;
    for (uint64_t t = (uint64_t)blockIdx.x * kWpb + threadIdx.x / kWave; t * kWave < total; t += nw) {
        const LbItem it = LbLocate(offs, np, t * kWave, total);
        uint32_t m = 0, v = 0, ui = 0, vj = 0;
        if (it.valid) {
            const uint32_t p = Y[it.p];
            v = p >> 4;
            ui = p & 15u;
            vj = G.pool[G.rec[v].off + it.j];
            const uint32_t s = S.vstate[vj];
            const uint32_t c = Code(s);
            m = c == kCodeNone ? 0u : (Q.uj_of[ui][c] & Lit(s));
        }
        const uint32_t per = both ? 2u : 1u;
        const uint32_t cnt = (uint32_t)__popc(m) * per;
        uint64_t at = WaveReserve(cnt, n_keys);
        if (cnt) {
            if (at + cnt > cap) {
                atomicOr(ovf, 1ull);
            } else {
                const uint64_t nu = S.new_of[v], nv = S.new_of[vj];
                while (m) {
                    const uint32_t uj = __ffs(m) - 1;
                    m &= m - 1;
                    keys[at++] = ((uint64_t)Q.eidx[ui][uj] << Q.ushift) | (nu << 32) | nv;
                    if (both) keys[at++] = ((uint64_t)Q.eidx[uj][ui] << Q.ushift) | (nv << 32) | nu;
                }
            }
        }
    }
}

// ConstructUpdate: every pair (v1, v2) of the batch, lit u1 of v1 and lit
// neighbour u2 of u1 labelled like v2 -> update and delete keys for
// (v1 -> v2, (u1, u2)) and (v2 -> v1, (u2, u1)); directions that received a
// key are flagged for edge_ok.
__global__ void KS1EmitUpdates(QTab Q, SView S, const int32_t* __restrict__ ints, uint64_t start, uint64_t np,
                               uint32_t n, uint64_t* __restrict__ upd, uint64_t* __restrict__ del, uint64_t cap,
                               unsigned long long* __restrict__ n_upd, unsigned long long* __restrict__ ovf,
                               unsigned long long* __restrict__ dirmask) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long dm0 = 0, dm1 = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < np; base += stride) {
        const uint64_t i = base + threadIdx.x;
        uint32_t cnt = 0, v1 = 0, v2 = 0, s1 = 0, s2 = 0;
        if (i < np) {
            const int32_t a = ints[start + 2 * i], b = ints[start + 2 * i + 1];
            if (a >= 0 && b >= 0 && (uint32_t)a < n && (uint32_t)b < n && a != b) {
                v1 = (uint32_t)a;
                v2 = (uint32_t)b;
                s1 = S.vstate[v1];
                s2 = S.vstate[v2];
                if (Code(s1) != kCodeNone && Code(s2) != kCodeNone) {
                    uint32_t l1 = Lit(s1);
                    while (l1) {
                        const uint32_t u1 = __ffs(l1) - 1;
                        l1 &= l1 - 1;
                        cnt += 2u * (uint32_t)__popc(Q.uj_of[u1][Code(s2)] & Lit(s2));
                    }
                }
            }
        }
        uint64_t at = WaveReserve(cnt, n_upd);
        if (cnt) {
            if (at + cnt > cap) {
                atomicOr(ovf, 1ull);
            } else {
                const uint64_t n1 = S.new_of[v1], n2 = S.new_of[v2];
                uint32_t l1 = Lit(s1);
                while (l1) {
                    const uint32_t u1 = __ffs(l1) - 1;
                    l1 &= l1 - 1;
                    uint32_t m = Q.uj_of[u1][Code(s2)] & Lit(s2);
                    while (m) {
                        const uint32_t u2 = __ffs(m) - 1;
                        m &= m - 1;
                        const uint32_t d12 = Q.eidx[u1][u2], d21 = Q.eidx[u2][u1];
                        const uint64_t k12 = ((uint64_t)d12 << Q.ushift) | (n1 << 32) | n2;
                        const uint64_t k21 = ((uint64_t)d21 << Q.ushift) | (n2 << 32) | n1;
                        upd[at] = k12;
                        del[at++] = k12;
                        upd[at] = k21;
                        del[at++] = k21;
                        if (d12 < 64) dm0 |= 1ull << d12;
                        else dm1 |= 1ull << (d12 - 64);
                        if (d21 < 64) dm0 |= 1ull << d21;
                        else dm1 |= 1ull << (d21 - 64);
                    }
                }
            }
        }
    }
    if (dm0) atomicOr(&dirmask[0], dm0);
    if (dm1) atomicOr(&dirmask[1], dm1);
}

__global__ void KS1ClearOpt(SView S, const uint32_t* __restrict__ Y, const unsigned long long* __restrict__ n_dev) {
    const uint64_t n = *n_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = Y[i];
        atomicAnd(&S.vstate[p >> 4], ~(1u << (16 + (p & 15u))));
    }
}

__global__ void KS1ExportLit(const uint32_t* __restrict__ vstate, uint32_t n, uint32_t* __restrict__ out) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = Lit(vstate[v]);
}

unsigned Grid(uint64_t n, unsigned per = 256, unsigned cap = 8192) {
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + per - 1) / per, cap));
}

}  // namespace
}  // namespace cgd

// ======================================================================== host side

using namespace cgapi;

// The data graph shared by the queries of a context (CaLiG's G.nei and
// labels, index/calig.h:28-33, index/calig.cpp:166-216), with the update
// stream resident in HBM and the current batch's new edges.
struct cgbdsm_graph {
    cgbdsm_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    uint32_t n = 0, vbits = 1;
    uint64_t m_dir = 0;            // directed adjacency entries
    DBuf<int32_t> labels;
    DBuf<MRec> rec;
    DBuf<uint32_t> pool;
    Counters* d_cnt = nullptr;
    Counters* h_cnt = nullptr;
    DBuf<int32_t> stream;          // update ints (two per pair; negative ids: deletions)
    uint64_t stream_ints = 0;
    uint64_t batches = 0;          // batches applied
    uint64_t b_start = 0, b_pairs = 0;  // the latest batch: ints [b_start, b_start + 2 * b_pairs)
    DBuf<uint64_t> newe;           // its new directed edges, sorted (src << vbits) | dst
    uint64_t n_new = 0;
    DBuf<uint64_t> dele, kd;       // its deleted directed edges (sorted, as newe); deletion key scratch
    uint64_t n_dele = 0;
    DBuf<uint64_t> k, alt, seg_start;
    DBuf<uint32_t> flags, seg_row, sc_val, sc_pos, hist;
    DBuf<uint8_t> scan;
    hipEvent_t ev = nullptr;       // latest batch applied on s
    std::mutex mu;
    std::vector<cgbdsm_query*> attached;
};

namespace cgd {

// Device Stage-1 state of one query.
struct S1Query {
    cgbdsm_graph* g = nullptr;
    QTab tab{};
    uint64_t batches = 0;          // graph batches consumed
    uint64_t pairs = 0;            // label-compatible pairs (list capacity)
    DBuf<uint32_t> vstate, queued, dead, new_of, old_of, newv, prefix;
    DBuf<uint32_t> L[6];           // work lists: 0/1 check lists, 2 died, 3 big, 4 touched/activated, 5 opt
    DBuf<uint32_t> L2[2];          // activation: touched list of the next round, activated list
    DBuf<uint16_t> died_bits;
    DBuf<uint64_t> offs;
    DBuf<uint64_t> kd;  // main-index entries a batch's deletions remove: [deleted edges | dead pairs]
    uint64_t n_dedge = 0;              // entries of the deleted edges (the negative update keys)
    uint8_t neg_ok[kMaxQE] = {};       // edge_ok of the negative pass
    DBuf<uint64_t> sv_ins, sv_del, sv_upd;  // the positive batch's key buffers during the negative pass
    S1Cnt* d_c = nullptr;
    S1Cnt* h_c = nullptr;
    uint32_t dv = 0;
    // statistics of the last call
    double ms = 0;
    uint64_t rounds_act = 0, rounds_del = 0, activated = 0, killed = 0, new_pairs = 0;
};

}  // namespace cgd

namespace cgapi {
namespace {

template <class T>
void GEnsure(cgbdsm_graph* g, DBuf<T>& b, size_t n) {
    if (n <= b.cap) return;
    const size_t nc = std::max<size_t>({n, b.cap + b.cap / 2, 1024});
    T* np = (T*)DevAlloc(g->ctx, nc * sizeof(T));
    if (b.p) {
        CGD_HIP(hipStreamSynchronize(g->s));
        DevFree(g->ctx, b.p, b.cap * sizeof(T));
    }
    b.p = np;
    b.cap = nc;
}

template <class T>
void GRelease(cgbdsm_ctx* ctx, DBuf<T>& b) {
    DevFree(ctx, b.p, b.cap * sizeof(T));
    b.p = nullptr;
    b.cap = 0;
}

void GReadCounters(cgbdsm_graph* g) {
    CGD_HIP(hipMemcpyAsync(g->h_cnt, g->d_cnt, sizeof(Counters), hipMemcpyDeviceToHost, g->s));
    CGD_HIP(hipStreamSynchronize(g->s));
}

Tables GraphTables(cgbdsm_graph* g) {
    Tables T{};
    T.m_rec = g->rec.p;
    T.pool = g->pool.p;
    T.ndir = 1;
    T.dv = g->n;
    return T;
}

GView View(cgbdsm_graph* g) { return GView{g->rec.p, g->pool.p, g->n}; }

void GKeyScratch(cgbdsm_graph* g, uint64_t n) {
    n = std::max<uint64_t>(n, 1);
    GEnsure(g, g->alt, n);
    GEnsure(g, g->flags, n + 1);
    GEnsure(g, g->seg_row, n + 1);
    GEnsure(g, g->seg_start, n + 2);
    GEnsure(g, g->sc_val, n);
    GEnsure(g, g->sc_pos, n);
    GEnsure(g, g->hist, RadixHistWords(n));
    GEnsure(g, g->scan, ScanScratchBytes(n + 2) + 64);
}

// Adjacency lists from n sorted unique keys in g->k (initial build).
void GraphLayout(cgbdsm_graph* g, uint64_t nk) {
    if (!nk) {
        FillU64((uint64_t*)&g->d_cnt->pool_top, 0, 1, g->s);
        if (!g->pool.p) {
            g->pool.p = (uint32_t*)DevAlloc(g->ctx, (1ull << 22) * sizeof(uint32_t));
            g->pool.cap = 1ull << 22;
        }
        return;
    }
    unsigned long long* n_u = &g->d_cnt->n_unique;
    unsigned long long* n_s = &g->d_cnt->n_segs;
    LaunchSegments(g->k.p, n_u, nk, g->vbits, g->flags.p, g->seg_row.p, g->seg_start.p, n_s, g->scan.p, g->s);
    GEnsure(g, g->alt, nk + 1);
    LaunchInitCaps(g->seg_start.p, n_s, nk, g->alt.p, g->s);
    ScanU64(g->alt.p, g->alt.p, std::max<uint64_t>(nk, 1), (uint64_t*)&g->d_cnt->pool_top, g->scan.p, g->s);
    GReadCounters(g);
    const uint64_t total = g->h_cnt->pool_top;
    const uint64_t want = std::max<uint64_t>(total + total / 2 + (16ull << 20), 1ull << 22);
    if (g->pool.cap < want) {
        GRelease(g->ctx, g->pool);
        g->pool.p = (uint32_t*)DevAlloc(g->ctx, want * sizeof(uint32_t));
        g->pool.cap = want;
    }
    LaunchInitLayout(GraphTables(g), g->k.p, n_u, nk, g->vbits, g->flags.p, g->seg_row.p, g->seg_start.p, n_s,
                     g->alt.p, g->d_cnt, g->s);
}

// Removes the sorted deleted-edge keys (all present) from the adjacency lists.
void GraphRemove(cgbdsm_graph* g, const uint64_t* keys, uint64_t nk) {
    if (!nk) return;
    unsigned long long* n_s = &g->d_cnt->n_segs;
    FillU64((uint64_t*)&g->d_cnt->n_unique, nk, 1, g->s);
    LaunchSegments(keys, &g->d_cnt->n_unique, nk, g->vbits, g->flags.p, g->seg_row.p, g->seg_start.p, n_s, g->scan.p,
                   g->s);
    FillU64((uint64_t*)&g->d_cnt->n_long, 0, 1, g->s);
    LaunchMergeRemove(GraphTables(g), keys, g->vbits, g->seg_row.p, g->seg_start.p, n_s, nk, g->flags.p, g->d_cnt,
                      g->s);
}

// Inserts the sorted new-edge keys into the adjacency lists (pool grows and
// the merge is relaunched on overflow; merged lists see nothing left).
void GraphMerge(cgbdsm_graph* g, const uint64_t* keys, uint64_t nk) {
    if (!nk) return;
    unsigned long long* n_s = &g->d_cnt->n_segs;
    FillU64((uint64_t*)&g->d_cnt->n_unique, nk, 1, g->s);
    LaunchSegments(keys, &g->d_cnt->n_unique, nk, g->vbits, g->flags.p, g->seg_row.p, g->seg_start.p, n_s, g->scan.p,
                   g->s);
    for (int attempt = 0;; attempt++) {
        FillU64((uint64_t*)&g->d_cnt->flags, 0, 1, g->s);
        FillU64((uint64_t*)&g->d_cnt->n_long, 0, 1, g->s);
        LaunchMergeInsert(GraphTables(g), keys, g->vbits, g->seg_row.p, g->seg_start.p, n_s, nk, g->sc_val.p,
                          g->sc_pos.p, g->flags.p, g->pool.cap, g->d_cnt, g->s);
        GReadCounters(g);
        if (!(g->h_cnt->flags & kFlagPool)) return;
        if (attempt >= 4) Throw(CGBDSM_ERR_OOM, "graph adjacency pool overflow after growth");
        const uint64_t used = g->h_cnt->pool_top;
        const uint64_t nc = std::max<uint64_t>(g->pool.cap * 2, used + (64ull << 20));
        uint32_t* np = (uint32_t*)DevAlloc(g->ctx, nc * sizeof(uint32_t));
        CGD_HIP(hipMemcpyAsync(np, g->pool.p, std::min<uint64_t>(used, g->pool.cap) * 4, hipMemcpyDeviceToDevice, g->s));
        CGD_HIP(hipStreamSynchronize(g->s));
        DevFree(g->ctx, g->pool.p, g->pool.cap * 4);
        g->pool.p = np;
        g->pool.cap = nc;
    }
}

// ---------------------------------------------------------------- per-query helpers

SView SV(S1Query* S) { return SView{S->vstate.p, S->queued.p, S->dead.p, S->new_of.p, S->old_of.p}; }

void S1Read(cgbdsm_query* q) {
    CGD_HIP(hipMemcpyAsync(q->s1->h_c, q->s1->d_c, sizeof(S1Cnt), hipMemcpyDeviceToHost, q->s));
    CGD_HIP(hipStreamSynchronize(q->s));
}

void S1Set(cgbdsm_query* q, unsigned long long S1Cnt::*f, uint64_t v) {
    FillU64((uint64_t*)&(q->s1->d_c->*f), v, 1, q->s);
}

// Offsets of the adjacency items of list L[0..n): offs[n] = total items.
// With n_dev, the list holds *n_dev <= n pairs (the rest count as empty).
void S1Offsets(cgbdsm_query* q, const uint32_t* L, uint64_t n, const unsigned long long* n_dev = nullptr) {
    S1Query* S = q->s1;
    Ensure(q, S->offs, n + 1);
    Ensure(q, q->scan_scratch, ScanScratchBytes(n + 2) + 64);
    if (n) KS1Deg<<<Grid(n), 256, 0, q->s>>>(View(S->g), L, n, n_dev, S->offs.p);
    ScanU64(S->offs.p, S->offs.p, n, S->offs.p + n, q->scan_scratch.p, q->s);
}

unsigned LbGrid(cgbdsm_query* q) { return (unsigned)q->ctx->num_cus * 8; }

// Deletion rounds from list L[src] (n pairs, all lit): check, kill the pairs
// that fail, queue the pairs their deaths may break, repeat until a round
// kills nothing.  opt_only: only pairs lit in this batch are re-checked.
// killed_list: list (and mark) the deaths of pairs lit before the batch.
void S1DeletionRounds(cgbdsm_query* q, int src, uint64_t n, bool opt_only, uint32_t* killed_list = nullptr) {
    S1Query* S = q->s1;
    const GView G = View(S->g);
    const SView V = SV(S);
    int cur = src;
    uint64_t rounds = 0, killed = 0;
    while (n) {
        unsigned long long* n_cur = &S->d_c->n_list[cur];
        S1Set(q, &S1Cnt::n_died, 0);
        S1Set(q, &S1Cnt::n_big, 0);
        KS1Check<<<Grid(n), kThreads, 0, q->s>>>(G, S->tab, V, S->L[cur].p, n_cur, S->L[2].p, S->died_bits.p,
                                                 &S->d_c->n_died, S->L[3].p, &S->d_c->n_big);
        KS1CheckWave<<<Grid(n, kWpb, 16384), kThreads, 0, q->s>>>(G, S->tab, V, S->L[3].p, &S->d_c->n_big,
                                                                 S->L[2].p, S->died_bits.p, &S->d_c->n_died);
        // The deaths (at most n) are applied and their neighbours queued with
        // the device count: one host round trip per round.
        KS1Apply<<<Grid(n), 256, 0, q->s>>>(S->tab, V, S->L[2].p, S->died_bits.p, &S->d_c->n_died,
                                            killed_list, &S->d_c->n_killed);
        S1Offsets(q, S->L[2].p, n, &S->d_c->n_died);
        const int nxt = cur == 0 ? 1 : 0;
        FillU64((uint64_t*)&S->d_c->n_list[nxt], 0, 1, q->s);
        KS1Expand<<<LbGrid(q), kThreads, 0, q->s>>>(G, S->tab, V, S->L[2].p, S->offs.p, n, opt_only ? 1u : 0u,
                                                   S->L[nxt].p, &S->d_c->n_list[nxt]);
        S1Read(q);
        rounds++;
        const uint64_t nd = S->h_c->n_died;
        if (!nd) break;
        killed += nd;
        n = S->h_c->n_list[nxt];
        cur = nxt;
    }
    S->rounds_del += rounds;
    S->killed += killed;
}

// Pairs of list `src` (n) -> keys in q->k_ins (both: entries of both ends).
// Returns the number of keys; grows the buffer and re-emits on overflow.
uint64_t S1EmitPairs(cgbdsm_query* q, const uint32_t* Y, uint64_t n, bool both) {
    S1Query* S = q->s1;
    S1Offsets(q, Y, n);
    for (int attempt = 0;; attempt++) {
        S1Set(q, &S1Cnt::n_keys, 0);
        S1Set(q, &S1Cnt::ovf, 0);
        if (n)
            KS1EmitPairs<<<LbGrid(q), kThreads, 0, q->s>>>(View(S->g), S->tab, SV(S), Y, S->offs.p, n, both ? 1u : 0u,
                                                          q->k_ins.p, q->k_ins.cap, &S->d_c->n_keys, &S->d_c->ovf);
        S1Read(q);
        const uint64_t nk = S->h_c->n_keys;
        if (!S->h_c->ovf) return nk;
        if (attempt) Throw(CGBDSM_ERR_DEVICE, "index key emission overflowed twice");
        Ensure(q, q->k_ins, nk + 1024);
    }
}

// Main-index entries over the graph's deleted edges (lit set before the
// batch) -> S->kd[0..); returns their number.
uint64_t S1EmitDelEdges(cgbdsm_query* q) {
    S1Query* S = q->s1;
    const cgbdsm_graph* g = S->g;
    Ensure(q, S->kd, std::max<uint64_t>(S->kd.cap, 1024));
    for (int attempt = 0;; attempt++) {
        S1Set(q, &S1Cnt::n_keys, 0);
        S1Set(q, &S1Cnt::ovf, 0);
        FillU64((uint64_t*)&S->d_c->dirmask[0], 0, 2, q->s);
        KS1DelEdgeKeys<<<Grid(g->n_dele), 256, 0, q->s>>>(S->tab, SV(S), g->dele.p, g->n_dele, g->vbits, S->kd.p,
                                                          S->kd.cap, &S->d_c->n_keys, &S->d_c->ovf,
                                                          &S->d_c->dirmask[0]);
        S1Read(q);
        const uint64_t nk = S->h_c->n_keys;
        if (!S->h_c->ovf) {
            auto on = [&](uint32_t d) { return (S->h_c->dirmask[d >> 6] >> (d & 63)) & 1ull; };
            for (uint32_t e = 0; e < q->qe; e++) S->neg_ok[e] = on(q->hplan.sa[e]) && on(q->hplan.sb[e]) ? 1 : 0;
            return nk;
        }
        if (attempt) Throw(CGBDSM_ERR_DEVICE, "deleted-edge key emission overflowed twice");
        Ensure(q, S->kd, nk + 1024);
    }
}

// Entries of the nk killed pairs in L2[1] -> S->kd[base..); clears their
// marks; returns the new end of S->kd.
uint64_t S1EmitKilled(cgbdsm_query* q, uint64_t base, uint64_t nk) {
    S1Query* S = q->s1;
    S1Offsets(q, S->L2[1].p, nk);
    uint64_t end = base;
    for (int attempt = 0;; attempt++) {
        FillU64((uint64_t*)&S->d_c->n_keys, base, 1, q->s);
        S1Set(q, &S1Cnt::ovf, 0);
        KS1EmitKilled<<<LbGrid(q), kThreads, 0, q->s>>>(View(S->g), S->tab, SV(S), S->L2[1].p, S->offs.p, nk, S->kd.p,
                                                       S->kd.cap, &S->d_c->n_keys, &S->d_c->ovf);
        S1Read(q);
        end = S->h_c->n_keys;
        if (!S->h_c->ovf) break;
        if (attempt) Throw(CGBDSM_ERR_DEVICE, "killed-pair key emission overflowed twice");
        Ensure(q, S->kd, end + 1024, true, base);
    }
    KS1ClearKilled<<<Grid(nk), 256, 0, q->s>>>(SV(S), S->L2[1].p, &S->d_c->n_killed);
    return end;
}

// New compressed ids for the vertices flagged in S->newv; returns the new dv.
uint32_t S1AssignIds(cgbdsm_query* q) {
    S1Query* S = q->s1;
    const uint64_t nw = (S->g->n + 31) / 32;
    Ensure(q, q->scan_scratch, ScanScratchBytes(nw + 2) + 64);
    KPopc<<<Grid(nw), 256, 0, q->s>>>(S->newv.p, nw, S->prefix.p);
    ScanU32(S->prefix.p, S->prefix.p, nw, (uint64_t*)&S->d_c->newv, q->scan_scratch.p, q->s);
    S1Read(q);
    const uint64_t dv = (uint64_t)S->dv + S->h_c->newv;
    if (dv > (1ull << (32 - q->ebits)) || dv > q->hplan.v_mask)
        Throw(CGBDSM_ERR_LIMIT, "compressed vertex count exceeds the key / row-tag width");
    Ensure(q, S->old_of, std::max<uint64_t>(dv, 1), true, S->dv);
    KS1AssignIds<<<Grid(nw), 256, 0, q->s>>>(S->newv.p, S->prefix.p, nw, S->dv, SV(S));
    return (uint32_t)dv;
}

QTab MakeQTab(const cgbdsm_query* q, const int32_t* qlabels) {
    QTab Q{};
    const cgbdsm_plan& P = q->plan;
    Q.qv = P.qv;
    for (uint32_t u = 0; u < P.qv; u++) {
        uint32_t c = Q.ndl;
        for (uint32_t k = 0; k < Q.ndl; k++)
            if (Q.dlabel[k] == qlabels[u]) c = k;
        if (c == Q.ndl) Q.dlabel[Q.ndl++] = qlabels[u];
        Q.qcode[u] = (uint8_t)c;
        Q.by_code[c] |= (uint16_t)(1u << u);
    }
    for (uint32_t a = 0; a < (uint32_t)kMaxQV; a++)
        for (uint32_t b = 0; b < (uint32_t)kMaxQV; b++) Q.eidx[a][b] = 0xFF;
    for (uint32_t a = 0; a < P.qv; a++)
        for (uint32_t b = 0; b < P.qv; b++) {
            const uint8_t e = P.eidx[a * kMaxQV + b];
            Q.eidx[a][b] = e;
            if (e != CGBDSM_NO_EDGE) Q.nbr[a] |= (uint16_t)(1u << b);
        }
    for (uint32_t u = 0; u < P.qv; u++) {
        for (uint32_t c = 0; c < 16; c++) Q.uj_of[u][c] = c < Q.ndl ? (uint16_t)(Q.nbr[u] & Q.by_code[c]) : 0;
        for (uint32_t c = 0; c < Q.ndl; c++) {
            const uint32_t m = Q.uj_of[u][c];
            if (__builtin_popcount(m) < 2) continue;
            Q.grp[u] |= (uint16_t)m;
            for (uint32_t uj = 0; uj < P.qv; uj++)
                if ((m >> uj) & 1u) Q.gneed[u] |= (uint64_t)__builtin_popcount(m) << (4 * uj);
        }
    }
    Q.ushift = q->ushift;
    return Q;
}

void S1Free(cgbdsm_query* q) {
    S1Query* S = q->s1;
    if (!S) return;
    for (auto* b : {&S->vstate, &S->queued, &S->dead, &S->new_of, &S->old_of, &S->newv, &S->prefix, &S->L[0], &S->L[1],
                    &S->L[2], &S->L[3], &S->L[4], &S->L[5], &S->L2[0], &S->L2[1]})
        Release(q, *b);
    Release(q, S->died_bits);
    Release(q, S->offs);
    for (auto* b : {&S->kd, &S->sv_ins, &S->sv_del, &S->sv_upd}) Release(q, *b);
    if (S->d_c) DevFree(q->ctx, S->d_c, sizeof(S1Cnt));
    if (S->h_c) (void)hipHostFree(S->h_c);
    delete S;
    q->s1 = nullptr;
}

}  // namespace

// Called by cgbdsm_query_destroy.
void S1Detach(cgbdsm_query* q) {
    if (!q->s1) return;
    cgbdsm_graph* g = q->s1->g;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        g->attached.erase(std::remove(g->attached.begin(), g->attached.end(), q), g->attached.end());
    }
    S1Free(q);
}

}  // namespace cgapi

extern "C" {

cgbdsm_status cgbdsm_graph_create(cgbdsm_ctx* ctx, uint32_t n, const int32_t* labels, const uint32_t* edges,
                                  uint64_t m, cgbdsm_graph** out) {
    return Guard([&] {
        if (!ctx || !out || (n && !labels) || (m && !edges)) Throw(CGBDSM_ERR_INVALID, "null argument");
        if (n == 0 || n >= (1u << 28)) Throw(CGBDSM_ERR_LIMIT, "data graph needs 1 .. 2^28 - 1 vertices");
        CGD_HIP(hipSetDevice(ctx->device));
        auto* g = new cgbdsm_graph;
        try {
            g->ctx = ctx;
            g->n = n;
            g->vbits = BitsFor(n);
            CGD_HIP(hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking));
            CGD_HIP(hipEventCreateWithFlags(&g->ev, hipEventDisableTiming));
            g->d_cnt = (Counters*)DevAlloc(ctx, sizeof(Counters));
            CGD_HIP(hipHostMalloc((void**)&g->h_cnt, sizeof(Counters), hipHostMallocDefault));
            FillU64((uint64_t*)g->d_cnt, 0, sizeof(Counters) / 8, g->s);
            GEnsure(g, g->labels, n);
            CGD_HIP(hipMemcpyAsync(g->labels.p, labels, (size_t)n * 4, hipMemcpyHostToDevice, g->s));
            GEnsure(g, g->rec, n);
            FillU64((uint64_t*)g->rec.p, 0, (uint64_t)n * 2, g->s);
            const uint64_t nk = 2 * m;
            GEnsure(g, g->k, std::max<uint64_t>(nk, 1));
            GKeyScratch(g, nk);
            DBuf<uint32_t> e;
            e.p = (uint32_t*)DevAlloc(ctx, std::max<uint64_t>(m, 1) * 8);
            e.cap = std::max<uint64_t>(m, 1) * 2;
            if (m) CGD_HIP(hipMemcpyAsync(e.p, edges, m * 8, hipMemcpyHostToDevice, g->s));
            if (m) KEdgeKeys<<<Grid(m), 256, 0, g->s>>>(e.p, m, n, g->vbits, g->k.p);
            RadixSortU64(g->k.p, g->alt.p, nk, 2 * g->vbits + 1, g->hist.p, g->s);
            // invalid pairs (~0) sort last; keep the valid prefix, unique
            LaunchUnique(g->k.p, g->alt.p, nk, g->flags.p, g->scan.p, &g->d_cnt->n_unique, g->s);
            GReadCounters(g);
            uint64_t nu = g->h_cnt->n_unique;
            CGD_HIP(hipStreamSynchronize(g->s));
            DevFree(ctx, e.p, std::max<uint64_t>(m, 1) * 8);
            if (nu) {
                uint64_t last = 0;
                CGD_HIP(hipMemcpy(&last, g->alt.p + nu - 1, 8, hipMemcpyDeviceToHost));
                if (last == ~0ull) nu--;
            }
            CGD_HIP(hipMemcpyAsync(g->k.p, g->alt.p, std::max<uint64_t>(nu, 1) * 8, hipMemcpyDeviceToDevice, g->s));
            FillU64((uint64_t*)&g->d_cnt->n_unique, nu, 1, g->s);
            GraphLayout(g, nu);
            g->m_dir = nu;
            CGD_HIP(hipStreamSynchronize(g->s));
        } catch (...) {
            cgbdsm_graph_destroy(g);
            throw;
        }
        *out = g;
    });
}

cgbdsm_status cgbdsm_graph_destroy(cgbdsm_graph* g) {
    return Guard([&] {
        if (!g) return;
        {
            std::lock_guard<std::mutex> lk(g->mu);
            if (!g->attached.empty()) Throw(CGBDSM_ERR_INVALID, "queries are still attached to the graph");
        }
        if (g->s) (void)hipStreamSynchronize(g->s);
        cgbdsm_ctx* c = g->ctx;
        GRelease(c, g->labels);
        GRelease(c, g->rec);
        GRelease(c, g->pool);
        GRelease(c, g->stream);
        GRelease(c, g->newe);
        for (auto* b : {&g->k, &g->alt, &g->seg_start, &g->dele, &g->kd}) GRelease(c, *b);
        for (auto* b : {&g->flags, &g->seg_row, &g->sc_val, &g->sc_pos, &g->hist}) GRelease(c, *b);
        GRelease(c, g->scan);
        if (g->d_cnt) DevFree(c, g->d_cnt, sizeof(Counters));
        if (g->h_cnt) (void)hipHostFree(g->h_cnt);
        if (g->ev) (void)hipEventDestroy(g->ev);
        if (g->s) (void)hipStreamDestroy(g->s);
        delete g;
    });
}

cgbdsm_status cgbdsm_graph_set_stream(cgbdsm_graph* g, const int32_t* ints, uint64_t n_ints) {
    return Guard([&] {
        if (!g || (n_ints && !ints)) Throw(CGBDSM_ERR_INVALID, "null argument");
        GEnsure(g, g->stream, std::max<uint64_t>(n_ints, 1));
        if (n_ints) CGD_HIP(hipMemcpyAsync(g->stream.p, ints, n_ints * 4, hipMemcpyHostToDevice, g->s));
        CGD_HIP(hipStreamSynchronize(g->s));
        g->stream_ints = n_ints;
    });
}

cgbdsm_status cgbdsm_graph_apply_range(cgbdsm_graph* g, uint64_t start, uint64_t count) {
    return Guard([&] {
        if (!g) Throw(CGBDSM_ERR_INVALID, "null graph");
        if (start > g->stream_ints) Throw(CGBDSM_ERR_INVALID, "batch starts beyond the stream");
        {
            std::lock_guard<std::mutex> lk(g->mu);
            for (cgbdsm_query* q : g->attached)
                if (q->s1->batches != g->batches)
                    Throw(CGBDSM_ERR_INVALID, "an attached query has not consumed the previous batch (cgbdsm_calig_update)");
        }
        const uint64_t end = std::min<uint64_t>(start + count, g->stream_ints);
        const uint64_t np = (end - start) / 2;  // main_calig.cpp:220-222 / calig.cpp:973-976
        const uint64_t nk = 2 * np;
        GEnsure(g, g->k, std::max<uint64_t>(nk, 1));
        GEnsure(g, g->kd, std::max<uint64_t>(nk, 1));
        GKeyScratch(g, nk);
        GEnsure(g, g->newe, std::max<uint64_t>(nk, 1));
        GEnsure(g, g->dele, std::max<uint64_t>(nk, 1));
        uint64_t n_new = 0, n_dele = 0;
        if (np) {
            KGraphPairKeys<<<Grid(np), 256, 0, g->s>>>(g->stream.p, start, np, g->n, g->vbits, g->k.p, g->kd.p);
            // Both kinds are judged against the graph before the batch
            // (calig.cpp:980-998): insertions of absent edges, deletions of
            // present ones; so no edge is both.
            const uint32_t kb = 2 * g->vbits + 1;
            RadixSortU64(g->k.p, g->alt.p, nk, kb, g->hist.p, g->s);
            LaunchUnique(g->k.p, g->alt.p, nk, g->flags.p, g->scan.p, &g->d_cnt->n_unique, g->s);
            KGraphAbsent<<<Grid(nk), 256, 0, g->s>>>(g->alt.p, &g->d_cnt->n_unique, nk, g->vbits, View(g), 0u,
                                                     g->sc_val.p);
            ScanU32(g->sc_val.p, g->sc_pos.p, nk, (uint64_t*)&g->d_cnt->n_segs, g->scan.p, g->s);
            KCompactKeys<<<Grid(nk), 256, 0, g->s>>>(g->alt.p, nk, g->sc_val.p, g->sc_pos.p, g->newe.p);
            RadixSortU64(g->kd.p, g->alt.p, nk, kb, g->hist.p, g->s);
            LaunchUnique(g->kd.p, g->alt.p, nk, g->flags.p, g->scan.p, &g->d_cnt->n_unique, g->s);
            KGraphAbsent<<<Grid(nk), 256, 0, g->s>>>(g->alt.p, &g->d_cnt->n_unique, nk, g->vbits, View(g), 1u,
                                                     g->sc_val.p);
            ScanU32(g->sc_val.p, g->sc_pos.p, nk, (uint64_t*)&g->d_cnt->u_rows, g->scan.p, g->s);
            KCompactKeys<<<Grid(nk), 256, 0, g->s>>>(g->alt.p, nk, g->sc_val.p, g->sc_pos.p, g->dele.p);
            GReadCounters(g);
            n_new = g->h_cnt->n_segs;
            n_dele = g->h_cnt->u_rows;
            GraphRemove(g, g->dele.p, n_dele);
            GraphMerge(g, g->newe.p, n_new);
        }
        g->n_new = n_new;
        g->n_dele = n_dele;
        g->m_dir = g->m_dir + n_new - n_dele;
        g->b_start = start;
        g->b_pairs = np;
        g->batches++;
        CGD_HIP(hipEventRecord(g->ev, g->s));
        CGD_HIP(hipStreamSynchronize(g->s));
    });
}

cgbdsm_status cgbdsm_graph_info(cgbdsm_graph* g, uint64_t* batches, uint64_t* directed_edges, uint64_t* last_new) {
    return Guard([&] {
        if (!g) Throw(CGBDSM_ERR_INVALID, "null graph");
        if (batches) *batches = g->batches;
        if (directed_edges) *directed_edges = g->m_dir;
        if (last_new) *last_new = g->n_new;
    });
}

cgbdsm_status cgbdsm_calig_attach(cgbdsm_query* q, cgbdsm_graph* g, const int32_t* qlabels, uint32_t* cardinalities,
                                  float* avg_degrees) {
    return Guard([&] {
        if (!q || !g || !qlabels) Throw(CGBDSM_ERR_INVALID, "null argument");
        if (q->ctx != g->ctx) Throw(CGBDSM_ERR_INVALID, "query and graph belong to different contexts");
        if (q->s1 || q->inited) Throw(CGBDSM_ERR_INVALID, "query already has an index");
        auto* S = new S1Query;
        q->s1 = S;
        try {
            S->g = g;
            S->tab = MakeQTab(q, qlabels);
            S->batches = g->batches;
            const uint32_t n = g->n;
            S->d_c = (S1Cnt*)DevAlloc(q->ctx, sizeof(S1Cnt));
            CGD_HIP(hipHostMalloc((void**)&S->h_c, sizeof(S1Cnt), hipHostMallocDefault));
            FillU64((uint64_t*)S->d_c, 0, sizeof(S1Cnt) / 8, q->s);
            CGD_HIP(hipStreamWaitEvent(q->s, g->ev, 0));
            Ensure(q, S->vstate, n);
            Ensure(q, S->queued, n);
            Ensure(q, S->new_of, n);
            Ensure(q, S->dead, ((uint64_t)n * S->tab.qv + 1) / 2);
            const uint64_t nw = (n + 31) / 32;
            Ensure(q, S->newv, nw + 1);
            Ensure(q, S->prefix, nw + 1);
            FillU32(S->queued.p, 0, n, q->s);
            FillU32(S->new_of.p, ~0u, n, q->s);
            FillU32(S->dead.p, 0, ((uint64_t)n * S->tab.qv + 1) / 2, q->s);
            FillU32(S->newv.p, 0, nw + 1, q->s);
            // constructCand: every label-compatible pair starts lit.
            KS1Codes<<<Grid(n), 256, 0, q->s>>>(g->labels.p, n, S->tab, S->vstate.p, 1u, &S->d_c->pairs);
            S1Read(q);
            S->pairs = S->h_c->pairs;
            const uint64_t cap = std::max<uint64_t>(S->pairs, 1);
            for (auto& l : S->L) Ensure(q, l, cap);
            for (auto& l : S->L2) Ensure(q, l, cap);
            Ensure(q, S->died_bits, cap);
            // staticFilter: deletion rounds from every pair.
            FillU64((uint64_t*)&S->d_c->n_list[0], 0, 1, q->s);
            KS1ListPairs<<<Grid(n), 256, 0, q->s>>>(S->vstate.p, n, S->tab, 0u, S->L[0].p, &S->d_c->n_list[0]);
            S1Read(q);
            S1DeletionRounds(q, 0, S->h_c->n_list[0], false);
            // ConvertGlobalIndex host half: ids in ascending original id,
            // then the keys of every lit pair.
            KS1LitBitmap<<<Grid(nw), 256, 0, q->s>>>(S->vstate.p, n, S->newv.p);
            S->dv = 0;
            const uint32_t dv = S1AssignIds(q);
            FillU64((uint64_t*)&S->d_c->n_list[4], 0, 1, q->s);
            KS1ListPairs<<<Grid(n), 256, 0, q->s>>>(S->vstate.p, n, S->tab, 1u, S->L[4].p, &S->d_c->n_list[4]);
            S1Read(q);
            const uint64_t nl = S->h_c->n_list[4];
            Ensure(q, q->k_ins, std::max<uint64_t>(q->k_ins.cap, 1024));
            const uint64_t nk = S1EmitPairs(q, S->L[4].p, nl, false);
            S->dv = dv;
            {
                std::lock_guard<std::mutex> lk(g->mu);
                g->attached.push_back(q);
            }
            IndexInitDevice(q, nk, std::max<uint32_t>(dv, 1), cardinalities, avg_degrees);
        } catch (...) {
            S1Detach(q);
            throw;
        }
    });
}

cgbdsm_status cgbdsm_calig_update(cgbdsm_query* q) {
    return Guard([&] {
        if (!q || !q->s1) Throw(CGBDSM_ERR_INVALID, "query has no device Stage 1 (cgbdsm_calig_attach)");
        if (!q->have_order) Throw(CGBDSM_ERR_INVALID, "cgbdsm_query_set_plan must run before the first batch");
        CheckCanStage(q);
        S1Query* S = q->s1;
        cgbdsm_graph* g = S->g;
        if (S->batches + 1 != g->batches)
            Throw(CGBDSM_ERR_INVALID, "the query must consume every graph batch in order (query at " +
                                          std::to_string(S->batches) + ", graph at " + std::to_string(g->batches) + ")");
        const auto t0 = std::chrono::steady_clock::now();
        CGD_HIP(hipStreamWaitEvent(q->s, g->ev, 0));
        const GView G = View(g);
        const SView V = SV(S);
        S->rounds_act = S->rounds_del = S->activated = S->killed = 0;
        // 1. Optimistic activation from the new edges.
        FillU64((uint64_t*)S->d_c, 0, sizeof(S1Cnt) / 8, q->s);
        if (g->n_new)
            KS1TouchNew<<<Grid(g->n_new), 256, 0, q->s>>>(S->tab, V, g->newe.p, g->n_new, g->vbits, S->L2[0].p,
                                                          &S->d_c->n_list[0]);
        for (;;) {
            // L2[0]: touched (count n_list[0]) -> L[4] activated (n_list[1]), L[5] all activated (n_list[2])
            FillU64((uint64_t*)&S->d_c->n_list[1], 0, 1, q->s);
            KS1Activate<<<Grid(std::max<uint64_t>(S->pairs, 1), 256, 2048), 256, 0, q->s>>>(
                S->tab, V, S->L2[0].p, &S->d_c->n_list[0], S->L[4].p, &S->d_c->n_list[1], S->L[5].p,
                &S->d_c->n_list[2]);
            S1Read(q);
            const uint64_t na = S->h_c->n_list[1];
            S->rounds_act++;
            if (!na) break;
            S->activated += na;
            S1Offsets(q, S->L[4].p, na);
            FillU64((uint64_t*)&S->d_c->n_list[0], 0, 1, q->s);
            KS1Push<<<LbGrid(q), kThreads, 0, q->s>>>(G, S->tab, V, S->L[4].p, S->offs.p, na, S->L2[0].p,
                                                     &S->d_c->n_list[0]);
        }
        const uint64_t n_opt = S->h_c->n_list[2];
        const uint64_t nd = g->n_dele;
        // 1b. Deleted edges: the main-index entries they carried (lit set
        // before the batch: lit and not activated in it).
        uint64_t n_kd = nd ? S1EmitDelEdges(q) : 0;
        S->n_dedge = n_kd;
        // 2. Deletion rounds over the optimistic pairs and, when edges were
        // deleted, the pairs from before the batch that lost one (then any
        // lit pair may die, so every lit neighbour of a death is re-checked).
        if (n_opt) CGD_HIP(hipMemcpyAsync(S->L[0].p, S->L[5].p, n_opt * 4, hipMemcpyDeviceToDevice, q->s));
        FillU64((uint64_t*)&S->d_c->n_list[0], n_opt, 1, q->s);
        uint64_t n_chk = n_opt;
        if (nd) {
            KS1SeedDeleted<<<Grid(nd), 256, 0, q->s>>>(S->tab, V, g->dele.p, nd, g->vbits, S->L[0].p,
                                                       &S->d_c->n_list[0]);
            S1Set(q, &S1Cnt::n_killed, 0);
            S1Read(q);
            n_chk = S->h_c->n_list[0];
        }
        if (n_chk) S1DeletionRounds(q, 0, n_chk, nd == 0, nd ? S->L2[1].p : nullptr);
        if (nd) {
            const uint64_t nk = S->h_c->n_killed;
            if (nk) n_kd = S1EmitKilled(q, n_kd, nk);
        }
        // 3. Survivors, new compressed ids, keys.
        FillU64((uint64_t*)&S->d_c->n_list[3], 0, 1, q->s);
        FillU64((uint64_t*)&S->d_c->n_list[2], n_opt, 1, q->s);
        if (n_opt)
            KS1Survivors<<<Grid(n_opt), 256, 0, q->s>>>(V, S->L[5].p, &S->d_c->n_list[2], S->L[4].p,
                                                        &S->d_c->n_list[3], S->newv.p);
        const uint32_t dv = S1AssignIds(q);  // reads the counters: n_list[3] = survivors
        const uint64_t ny = S->h_c->n_list[3];
        S->new_pairs = ny;
        const uint64_t n_ins = S1EmitPairs(q, S->L[4].p, ny, true);
        const uint64_t np = g->b_pairs;
        const uint64_t cap_u = std::max<uint64_t>({q->k_upd.cap, q->k_del.cap, 1024});
        Ensure(q, q->k_upd, cap_u);
        Ensure(q, q->k_del, cap_u);
        uint64_t n_upd = 0;
        for (int attempt = 0;; attempt++) {
            S1Set(q, &S1Cnt::n_upd, 0);
            S1Set(q, &S1Cnt::ovf, 0);
            FillU64((uint64_t*)&S->d_c->dirmask[0], 0, 2, q->s);
            if (np)
                KS1EmitUpdates<<<Grid(np), 256, 0, q->s>>>(S->tab, V, g->stream.p, g->b_start, np, g->n, q->k_upd.p,
                                                           q->k_del.p, std::min(q->k_upd.cap, q->k_del.cap),
                                                           &S->d_c->n_upd, &S->d_c->ovf, &S->d_c->dirmask[0]);
            S1Read(q);
            n_upd = S->h_c->n_upd;
            if (!S->h_c->ovf) break;
            if (attempt) Throw(CGBDSM_ERR_DEVICE, "update key emission overflowed twice");
            Ensure(q, q->k_upd, n_upd + 1024);
            Ensure(q, q->k_del, n_upd + 1024);
        }
        if (ny) KS1ClearOpt<<<Grid(ny), 256, 0, q->s>>>(V, S->L[4].p, &S->d_c->n_list[3]);
        if (n_kd) {
            // Del = the update keys (removed until the commit) + the entries
            // the deletions remove for good.
            Ensure(q, q->k_del, n_upd + n_kd, true, n_upd);
            CGD_HIP(hipMemcpyAsync(q->k_del.p + n_upd, S->kd.p, n_kd * 8, hipMemcpyDeviceToDevice, q->s));
        }
        uint8_t edge_ok[kMaxQE] = {};
        auto dir_on = [&](uint32_t d) { return (S->h_c->dirmask[d >> 6] >> (d & 63)) & 1ull; };
        for (uint32_t e = 0; e < q->qe; e++) edge_ok[e] = dir_on(q->hplan.sa[e]) && dir_on(q->hplan.sb[e]) ? 1 : 0;
        S->dv = dv;
        S->batches++;
        StageDeviceKeys(q, n_ins, n_upd + n_kd, n_upd, edge_ok, std::max<uint32_t>(dv, q->dv));
        CGD_HIP(hipStreamSynchronize(q->s));
        S->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    });
}

cgbdsm_status cgbdsm_calig_match_deleted(cgbdsm_query* q, unsigned long long* num_matches, uint64_t* per_edge) {
    return Guard([&] {
        if (!q || !q->s1) Throw(CGBDSM_ERR_INVALID, "query has no device Stage 1 (cgbdsm_calig_attach)");
        if (q->phase != 1 || q->s1->batches == 0)
            Throw(CGBDSM_ERR_INVALID, "cgbdsm_calig_match_deleted must follow cgbdsm_calig_update, before Stage 2");
        S1Query* S = q->s1;
        if (num_matches) *num_matches = 0;
        if (per_edge) memset(per_edge, 0, q->qe * sizeof(uint64_t));
        const uint64_t nneg = S->n_dedge;
        if (!nneg) return;
        auto chk = [](cgbdsm_status st) {
            if (st != CGBDSM_OK) Throw(st, g_err);
        };
        // Park the positive batch, stage the deleted edges' entries as both
        // the Del and the update keys (M loses them for good; they are never
        // committed back), run Stage 2, then restore the positive batch.
        const uint64_t n_ins = q->n_ins, n_del = q->n_del, n_upd = q->n_upd;
        const auto st_save = q->st;
        uint8_t ok_save[kMaxQE];
        memcpy(ok_save, q->edge_ok, sizeof(ok_save));
        std::swap(q->k_ins, S->sv_ins);
        std::swap(q->k_del, S->sv_del);
        std::swap(q->k_upd, S->sv_upd);
        auto restore = [&] {
            std::swap(q->k_ins, S->sv_ins);
            std::swap(q->k_del, S->sv_del);
            std::swap(q->k_upd, S->sv_upd);
            q->n_ins = n_ins;
            q->n_del = n_del;
            q->n_upd = n_upd;
            memcpy(q->edge_ok, ok_save, sizeof(ok_save));
            q->st.last_ins = st_save.last_ins;
            q->st.last_del = st_save.last_del;
            q->st.last_upd = st_save.last_upd;
            q->phase = 1;
        };
        try {
            StageDeviceKeys(q, 0, nneg, nneg, S->neg_ok, q->dv);
            CGD_HIP(hipMemcpyAsync(q->k_del.p, S->kd.p, nneg * 8, hipMemcpyDeviceToDevice, q->s));
            CGD_HIP(hipMemcpyAsync(q->k_upd.p, S->kd.p, nneg * 8, hipMemcpyDeviceToDevice, q->s));
            chk(cgbdsm_transfer_update_all(q));
            int ne = 0;
            chk(cgbdsm_transfer_update_index(q, &ne));
            if (ne) {
                chk(cgbdsm_build_local_index(q));
                chk(cgbdsm_matching(q, num_matches, per_edge));
            }
        } catch (...) {
            restore();
            throw;
        }
        restore();
        S->n_dedge = 0;  // counted; a second call finds nothing left to count
    });
}

cgbdsm_status cgbdsm_calig_export_lit(cgbdsm_query* q, uint32_t* li, uint32_t n) {
    return Guard([&] {
        if (!q || !q->s1 || !li) Throw(CGBDSM_ERR_INVALID, "null argument or no device Stage 1");
        if (n != q->s1->g->n) Throw(CGBDSM_ERR_INVALID, "n must be the graph's vertex count");
        uint32_t* d = (uint32_t*)DevAlloc(q->ctx, (size_t)n * 4);
        KS1ExportLit<<<Grid(n), 256, 0, q->s>>>(q->s1->vstate.p, n, d);
        CGD_HIP(hipMemcpyAsync(li, d, (size_t)n * 4, hipMemcpyDeviceToHost, q->s));
        CGD_HIP(hipStreamSynchronize(q->s));
        DevFree(q->ctx, d, (size_t)n * 4);
    });
}

cgbdsm_status cgbdsm_calig_old_ids(cgbdsm_query* q, uint32_t* ids, uint32_t cap, uint32_t* n_out) {
    return Guard([&] {
        if (!q || !q->s1) Throw(CGBDSM_ERR_INVALID, "null argument or no device Stage 1");
        const uint32_t dv = q->s1->dv;
        if (n_out) *n_out = dv;
        if (ids && cap && dv) {
            CGD_HIP(hipMemcpyAsync(ids, q->s1->old_of.p, (size_t)std::min(cap, dv) * 4, hipMemcpyDeviceToHost, q->s));
            CGD_HIP(hipStreamSynchronize(q->s));
        }
    });
}

cgbdsm_status cgbdsm_calig_stats(cgbdsm_query* q, double* ms, uint64_t* out6) {
    return Guard([&] {
        if (!q || !q->s1) Throw(CGBDSM_ERR_INVALID, "null argument or no device Stage 1");
        const S1Query* S = q->s1;
        if (ms) *ms = S->ms;
        if (out6) {
            out6[0] = S->rounds_act;
            out6[1] = S->rounds_del;
            out6[2] = S->activated;
            out6[3] = S->killed;
            out6[4] = S->new_pairs;
            out6[5] = S->dv;
        }
    });
}

}  // extern "C"
