// GPMA layout of the main index on gfx950: a packed-memory array (gap array)
// of index keys, maintained by batch insert with density-bounded window
// rebalancing and batch delete, and extracted into the flat per-row lists
// the matching kernels read.
//
// Reference behaviour: CGCSM_GPMA (index/gpma_wrapper.cu:19-116) over
// Multi_GPMA (gpma/gpma.cuh:36-111): leaves with density thresholds 0.92 at
// the leaf level and 0.84 at the root (gpma/gpma.cuh:55-58), a locate /
// sort / level-wise rebalance / large-rebalance update (update_gpma_stage1..4,
// gpma/gpma.cu:1125-1295), and a flat per-row index rebuilt for matching
// every batch (index/calig_helper_gpma.cu:383-414).  The large rebalance
// there uses dynamic parallelism; here every step is a flat grid-stride
// kernel and the window choice is a sum tree walked per dirty leaf.
//
// Layout: one array of nleaves * kLeaf u64 keys, key = (row << 27) | v
// (row = owner * ndir + dir), each leaf left-packed (sorted keys, then
// kEmpty).  lo[leaf] is the leaf's separator: a key belongs to the last leaf
// whose separator is <= the key; lo is non-decreasing and lo[0] = 0.
#include "engine.h"
#include "kernels.h"
#include "wave.h"

namespace cgd {

namespace {

constexpr uint64_t kEmpty = ~0ull;

__device__ __forceinline__ uint64_t UpperBoundU64(const uint64_t* a, uint64_t n, uint64_t x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t m = (lo + hi) >> 1;
        if (a[m] <= x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__device__ __forceinline__ uint64_t LowerBoundU64(const uint64_t* a, uint64_t n, uint64_t x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t m = (lo + hi) >> 1;
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__device__ __forceinline__ uint64_t LowerBoundU32(const uint32_t* a, uint64_t n, uint32_t x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t m = (lo + hi) >> 1;
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// Upper density bound of a window of 2^level leaves (linear from the leaf
// bound to the root bound over the tree height).
__device__ __forceinline__ uint64_t WindowCap(uint32_t level, uint32_t height) {
    const double t = height ? (double)level / (double)height : 1.0;
    const double tau = 0.92 - (0.92 - 0.84) * t;
    return (uint64_t)(tau * (double)((uint64_t)kPmaLeaf << level));
}

__device__ __forceinline__ uint64_t LevelOff(uint32_t nleaves, uint32_t level) {
    // Level l of the sum tree starts after levels 0..l-1 (nleaves >> k entries each).
    return 2ull * nleaves - (2ull * nleaves >> level);
}

__global__ void KPmaEncode(const uint64_t* __restrict__ in, uint64_t n, uint32_t vbits, uint64_t* __restrict__ out) {
    const uint64_t vmask = (1ull << vbits) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = in[i];
        out[i] = ((k >> vbits) << kPmaVBits) | (k & vmask);
    }
}

// Even spread of n sorted keys over the leaves (bulk load / rebuild).
__global__ void KPmaSpread(const uint64_t* __restrict__ keys, uint64_t n, uint64_t* __restrict__ slot,
                           uint32_t* __restrict__ cnt, uint64_t* __restrict__ lo, uint32_t nleaves) {
    for (uint32_t leaf = blockIdx.x; leaf < nleaves; leaf += gridDim.x) {
        const uint64_t s0 = (uint64_t)leaf * n / nleaves, s1 = (uint64_t)(leaf + 1) * n / nleaves;
        const uint32_t c = (uint32_t)(s1 - s0);
        for (uint32_t t = threadIdx.x; t < kPmaLeaf; t += blockDim.x)
            slot[(uint64_t)leaf * kPmaLeaf + t] = t < c ? keys[s0 + t] : kEmpty;
        if (threadIdx.x == 0) {
            cnt[leaf] = c;
            lo[leaf] = leaf == 0 ? 0ull : (s0 < n ? keys[s0] : kEmpty);
        }
    }
}

__global__ void KPmaRoute(const uint64_t* __restrict__ keys, uint64_t n, const uint64_t* __restrict__ lo,
                          uint32_t nleaves, uint32_t* __restrict__ leaf_of) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t u = UpperBoundU64(lo, nleaves, keys[i]);
        leaf_of[i] = (uint32_t)(u ? u - 1 : 0);
    }
}

// keep[i] = 1 when key i is not stored yet (inserts of present keys are dropped).
__global__ void KPmaAbsent(const uint64_t* __restrict__ keys, uint64_t n, const uint32_t* __restrict__ leaf_of,
                           const uint64_t* __restrict__ slot, const uint32_t* __restrict__ cnt,
                           uint32_t* __restrict__ keep) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t leaf = leaf_of[i];
        const uint64_t* L = slot + (uint64_t)leaf * kPmaLeaf;
        const uint32_t c = cnt[leaf];
        const uint64_t p = LowerBoundU64(L, c, keys[i]);
        keep[i] = (p < c && L[p] == keys[i]) ? 0u : 1u;
    }
}

__global__ void KPmaCompact(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ leaf_of, uint64_t n,
                            const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos,
                            uint64_t* __restrict__ out_keys, uint32_t* __restrict__ out_leaf) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (keep[i]) {
            out_keys[pos[i]] = keys[i];
            out_leaf[pos[i]] = leaf_of[i];
        }
}

__global__ void KPmaInsCount(const uint32_t* __restrict__ leaf_of, const unsigned long long* __restrict__ n_dev,
                             uint32_t* __restrict__ ins_cnt) {
    // Per run of equal leaves, its end is added and its start subtracted
    // (mod 2^32): two atomics per run instead of one per key on a hot leaf.
    // Each maximal run adds its length, so any order of leaf_of is counted
    // right; the routed keys are sorted, so a leaf is one run.
    const uint64_t n = *n_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t l = leaf_of[i];
        if (i == 0 || leaf_of[i - 1] != l) atomicAdd(&ins_cnt[l], (uint32_t)(0u - (uint32_t)i));
        if (i + 1 == n || leaf_of[i + 1] != l) atomicAdd(&ins_cnt[l], (uint32_t)(i + 1));
    }
}

__global__ void KPmaLevel0(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ ins_cnt, uint32_t nleaves,
                           uint64_t* __restrict__ S) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nleaves; i += gridDim.x * blockDim.x)
        S[i] = (uint64_t)cnt[i] + ins_cnt[i];
}

__global__ void KPmaLevelUp(uint64_t* __restrict__ S, uint32_t nleaves, uint32_t level) {
    const uint64_t* prev = S + LevelOff(nleaves, level - 1);
    uint64_t* cur = S + LevelOff(nleaves, level);
    const uint32_t nw = nleaves >> level;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += gridDim.x * blockDim.x)
        cur[w] = prev[2 * w] + prev[2 * w + 1];
}

// The upper levels of the sum tree (at most 2048 windows each) in one
// block, level after level: one launch instead of one per level.
__global__ void __launch_bounds__(1024) KPmaLevelsTop(uint64_t* __restrict__ S, uint32_t nleaves, uint32_t l0,
                                                      uint32_t height) {
    for (uint32_t l = l0; l <= height; l++) {
        const uint64_t* prev = S + LevelOff(nleaves, l - 1);
        uint64_t* cur = S + LevelOff(nleaves, l);
        const uint32_t nw = nleaves >> l;
        for (uint32_t w = threadIdx.x; w < nw; w += blockDim.x) cur[w] = prev[2 * w] + prev[2 * w + 1];
        __syncthreads();
    }
}

// Each leaf with inserts takes the smallest window around it that stays
// under its level's density bound; no window fits: the array must grow.
__global__ void KPmaChoose(const uint32_t* __restrict__ ins_cnt, const uint64_t* __restrict__ S, uint32_t nleaves,
                           uint32_t height, uint8_t* __restrict__ chosen, Counters* __restrict__ cnt) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nleaves; i += gridDim.x * blockDim.x) {
        if (!ins_cnt[i]) continue;
        bool placed = false;
        for (uint32_t l = 0; l <= height; l++) {
            const uint32_t w = i >> l;
            if (S[LevelOff(nleaves, l) + w] <= WindowCap(l, height)) {
                chosen[LevelOff(nleaves, l) + w] = 1;
                placed = true;
                break;
            }
        }
        if (!placed) atomicOr(&cnt->flags, kFlagPool);
    }
}

// Chosen windows without a chosen ancestor are rebalanced; their leaves
// point at them (win_of).
__global__ void KPmaTop(const uint8_t* __restrict__ chosen, uint32_t nleaves, uint32_t height,
                        uint32_t* __restrict__ wins, unsigned long long* __restrict__ nwin,
                        uint32_t* __restrict__ win_of, unsigned long long* __restrict__ level_seen) {
    const uint64_t total = 2ull * nleaves - 1;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (uint64_t)gridDim.x * blockDim.x) {
        if (!chosen[idx]) continue;
        uint32_t l = 0;
        while (LevelOff(nleaves, l + 1) <= idx) l++;
        const uint32_t w = (uint32_t)(idx - LevelOff(nleaves, l));
        bool top = true;
        for (uint32_t a = l + 1; a <= height && top; a++)
            if (chosen[LevelOff(nleaves, a) + (w >> (a - l))]) top = false;
        if (!top) continue;
        const unsigned long long id = atomicAdd(nwin, 1ull);
        atomicMax(level_seen, (unsigned long long)l + 1);
        wins[id] = (l << 27) | w;
        const uint32_t a0 = w << l, a1 = (w + 1) << l;
        for (uint32_t leaf = a0; leaf < a1; leaf++) win_of[leaf] = (uint32_t)id;
    }
}

__global__ void KPmaWinMeta(const uint32_t* __restrict__ wins, const unsigned long long* __restrict__ nwin,
                            const uint32_t* __restrict__ gpos, const uint32_t* __restrict__ ins_leaf,
                            const unsigned long long* __restrict__ n_ins_dev, PmaWin* __restrict__ meta) {
    const uint64_t nw = *nwin, n_ins = *n_ins_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t l = wins[i] >> 27, w = wins[i] & ((1u << 27) - 1u);
        PmaWin m;
        m.a = w << l;
        m.nl = 1u << l;
        m.E = gpos[m.a + m.nl] - gpos[m.a];
        m.ilo = LowerBoundU32(ins_leaf, n_ins, m.a);
        m.ihi = LowerBoundU32(ins_leaf, n_ins, m.a + m.nl);
        meta[i] = m;
    }
}

// Existing keys of every window, compacted in order into the window's own
// slot range of the shadow array.
__global__ void KPmaGather(const uint64_t* __restrict__ slot, const uint32_t* __restrict__ cnt,
                           const uint32_t* __restrict__ gpos, const uint32_t* __restrict__ win_of,
                           const PmaWin* __restrict__ meta, uint64_t* __restrict__ shadow, uint32_t nleaves) {
    const uint64_t total = (uint64_t)nleaves * kPmaLeaf;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t leaf = (uint32_t)(s / kPmaLeaf), t = (uint32_t)(s % kPmaLeaf);
        const uint32_t wi = win_of[leaf];
        if (wi == 0xFFFFFFFFu || t >= cnt[leaf]) continue;
        const PmaWin m = meta[wi];
        shadow[(uint64_t)m.a * kPmaLeaf + (gpos[leaf] - gpos[m.a]) + t] = slot[s];
    }
}

__device__ __forceinline__ void PlaceAt(uint64_t* slot, const PmaWin& m, uint64_t rank, uint64_t key) {
    const uint64_t T = m.E + (m.ihi - m.ilo), nl = m.nl;
    const uint64_t j = ((rank + 1) * nl - 1) / T;
    const uint64_t start = j * T / nl;
    slot[((uint64_t)m.a + j) * kPmaLeaf + (rank - start)] = key;
}

// Merged rank of every key of a window (existing and inserted sets are
// disjoint), then its place in the even spread over the window's leaves.
__global__ void KPmaPlace(const uint64_t* __restrict__ shadow, const uint64_t* __restrict__ ins_keys,
                          const uint32_t* __restrict__ ins_leaf, const unsigned long long* __restrict__ n_ins_dev,
                          const uint32_t* __restrict__ win_of, const PmaWin* __restrict__ meta,
                          uint64_t* __restrict__ slot, uint32_t nleaves) {
    const uint64_t total = (uint64_t)nleaves * kPmaLeaf;
    const uint64_t n_ins = *n_ins_dev;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total + n_ins; p += stride) {
        if (p < total) {
            const uint32_t wi = win_of[p / kPmaLeaf];
            if (wi == 0xFFFFFFFFu) continue;
            const PmaWin m = meta[wi];
            const uint64_t e = p - (uint64_t)m.a * kPmaLeaf;
            if (e >= m.E) continue;
            const uint64_t key = shadow[p];
            const uint64_t r = e + LowerBoundU64(ins_keys + m.ilo, m.ihi - m.ilo, key);
            PlaceAt(slot, m, r, key);
        } else {
            const uint64_t i = p - total;
            const uint32_t wi = win_of[ins_leaf[i]];
            if (wi == 0xFFFFFFFFu) continue;  // only after an overflow, which rebuilds instead
            const PmaWin m = meta[wi];
            const uint64_t key = ins_keys[i];
            const uint64_t r = (i - m.ilo) + LowerBoundU64(shadow + (uint64_t)m.a * kPmaLeaf, m.E, key);
            PlaceAt(slot, m, r, key);
        }
    }
}

// Leaf counts, empty tails and separators of rebalanced leaves.  The first
// leaf of a window keeps its separator; an empty leaf takes the next key.
__global__ void KPmaFinalize(const uint32_t* __restrict__ win_of, const PmaWin* __restrict__ meta,
                             uint64_t* __restrict__ slot, uint32_t* __restrict__ cnt, uint64_t* __restrict__ lo,
                             uint32_t nleaves) {
    for (uint32_t leaf = blockIdx.x; leaf < nleaves; leaf += gridDim.x) {
        const uint32_t wi = win_of[leaf];
        if (wi == 0xFFFFFFFFu) continue;
        const PmaWin m = meta[wi];
        const uint64_t T = m.E + (m.ihi - m.ilo), nl = m.nl, j = leaf - m.a;
        const uint64_t s0 = j * T / nl, s1 = (j + 1) * T / nl;
        const uint32_t c = (uint32_t)(s1 - s0);
        for (uint32_t t = c + threadIdx.x; t < kPmaLeaf; t += blockDim.x) slot[(uint64_t)leaf * kPmaLeaf + t] = kEmpty;
        if (threadIdx.x == 0) {
            cnt[leaf] = c;
            if (j > 0) {
                uint64_t sep;
                if (s0 < T) {
                    // the key of merged rank s0 (first key of this or a later leaf)
                    const uint64_t jj = ((s0 + 1) * nl - 1) / T;
                    sep = slot[((uint64_t)m.a + jj) * kPmaLeaf + (s0 - jj * T / nl)];
                } else {
                    const uint32_t b = m.a + m.nl;
                    sep = b < nleaves ? lo[b] : kEmpty;
                }
                lo[leaf] = sep;
            }
        }
    }
}

// Deletes: one wavefront per leaf that holds delete keys (a wave per sorted
// delete key, which goes on only when its key is the first of its leaf), so
// that scattered deletes compact their leaves in parallel rather than up to
// 64 leaves one after another per wave.
__global__ void __launch_bounds__(256) KPmaDelete(const uint64_t* __restrict__ dkeys, const uint32_t* __restrict__ dleaf,
                                                  uint64_t n, uint64_t* __restrict__ slot, uint32_t* __restrict__ cnt) {
    const uint32_t lane = LaneId();
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / kWave);
    for (uint64_t h = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; h < n; h += nw) {
        const uint32_t leaf = Uniform(dleaf[h]);
        if (h > 0 && Uniform(dleaf[h - 1]) == leaf) continue;
        uint64_t e = h;  // end of this leaf's delete keys
        {
            uint64_t lo = h, hi = n;
            while (lo < hi) {
                const uint64_t m = (lo + hi) >> 1;
                if (dleaf[m] <= leaf) lo = m + 1;
                else hi = m;
            }
            e = lo;
        }
        uint64_t* L = slot + (uint64_t)leaf * kPmaLeaf;
        const uint32_t c = Uniform(cnt[leaf]);
        uint32_t w = 0;
        for (uint32_t k = 0; k < c; k += kWave) {
            const uint32_t t = k + lane;
            const bool valid = t < c;
            const uint64_t key = valid ? L[t] : kEmpty;
            bool del = false;
            if (valid) {
                const uint64_t p = h + LowerBoundU64(dkeys + h, e - h, key);
                del = p < e && dkeys[p] == key;
            }
            const bool keep = valid && !del;
            const uint64_t m = __ballot(keep);
            WaveSync();
            if (keep) L[w + RankBelow(m)] = key;
            WaveSync();
            w += (uint32_t)__popcll(m);
        }
        for (uint32_t t = w + lane; t < c; t += kWave) L[t] = kEmpty;
        if (lane == 0) cnt[leaf] = w;
        WaveSync();
    }
}

// Extraction: keys in array order into a flat buffer (gpos = exclusive scan of cnt).
__global__ void KPmaFlat(const uint64_t* __restrict__ slot, const uint32_t* __restrict__ cnt,
                         const uint32_t* __restrict__ gpos, uint32_t nleaves, uint64_t* __restrict__ flat) {
    const uint64_t total = (uint64_t)nleaves * kPmaLeaf;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < total;
         s += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t leaf = (uint32_t)(s / kPmaLeaf), t = (uint32_t)(s % kPmaLeaf);
        if (t < cnt[leaf]) flat[gpos[leaf] + t] = slot[s];
    }
}

__global__ void KPmaClearRecs(MRec* __restrict__ m_rec, uint64_t rows) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * blockDim.x)
        m_rec[r] = MRec{0, 0, 0};
}

// Flat keys -> list values in the pool and one record per non-empty row
// (written by the thread holding the row's last key).
__global__ void KPmaRows(const uint64_t* __restrict__ flat, uint64_t n, uint32_t* __restrict__ pool,
                         MRec* __restrict__ m_rec) {
    const uint64_t vmask = (1ull << kPmaVBits) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t k = flat[i];
        pool[i] = (uint32_t)(k & vmask);
        const uint64_t row = k >> kPmaVBits;
        if (i + 1 == n || (flat[i + 1] >> kPmaVBits) != row) {
            const uint64_t start = LowerBoundU64(flat, i + 1, row << kPmaVBits);
            const uint32_t len = (uint32_t)(i + 1 - start);
            m_rec[row] = MRec{start, len, len};
        }
    }
}

unsigned GridFor64(uint64_t n, unsigned per = 256, unsigned cap = 16384) {
    const uint64_t g = (n + per - 1) / per;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

void LaunchPmaEncode(const uint64_t* in, uint64_t n, uint32_t vbits, uint64_t* out, hipStream_t s) {
    if (n) KPmaEncode<<<GridFor64(n), 256, 0, s>>>(in, n, vbits, out);
}
void LaunchPmaSpread(const uint64_t* keys, uint64_t n, uint64_t* slot, uint32_t* cnt, uint64_t* lo, uint32_t nleaves,
                     hipStream_t s) {
    KPmaSpread<<<std::min<uint32_t>(nleaves, 16384), 256, 0, s>>>(keys, n, slot, cnt, lo, nleaves);
}
void LaunchPmaRoute(const uint64_t* keys, uint64_t n, const uint64_t* lo, uint32_t nleaves, uint32_t* leaf_of,
                    hipStream_t s) {
    if (n) KPmaRoute<<<GridFor64(n), 256, 0, s>>>(keys, n, lo, nleaves, leaf_of);
}
void LaunchPmaAbsent(const uint64_t* keys, uint64_t n, const uint32_t* leaf_of, const uint64_t* slot,
                     const uint32_t* cnt, uint32_t* keep, hipStream_t s) {
    if (n) KPmaAbsent<<<GridFor64(n), 256, 0, s>>>(keys, n, leaf_of, slot, cnt, keep);
}
void LaunchPmaCompact(const uint64_t* keys, const uint32_t* leaf_of, uint64_t n, const uint32_t* keep,
                      const uint32_t* pos, uint64_t* out_keys, uint32_t* out_leaf, hipStream_t s) {
    if (n) KPmaCompact<<<GridFor64(n), 256, 0, s>>>(keys, leaf_of, n, keep, pos, out_keys, out_leaf);
}
void LaunchPmaWindows(const uint32_t* cnt, uint32_t* ins_cnt, const uint32_t* ins_leaf,
                      const unsigned long long* n_ins_dev, uint64_t n_upper, uint64_t* S, uint8_t* chosen,
                      uint32_t nleaves, uint32_t height, Counters* counters, hipStream_t s) {
    FillU32(ins_cnt, 0, nleaves, s);
    if (n_upper) KPmaInsCount<<<GridFor64(n_upper), 256, 0, s>>>(ins_leaf, n_ins_dev, ins_cnt);
    KPmaLevel0<<<GridFor64(nleaves), 256, 0, s>>>(cnt, ins_cnt, nleaves, S);
    uint32_t l = 1;
    for (; l <= height && (nleaves >> l) > 2048; l++)
        KPmaLevelUp<<<GridFor64(nleaves >> l), 256, 0, s>>>(S, nleaves, l);
    if (l <= height) KPmaLevelsTop<<<1, 1024, 0, s>>>(S, nleaves, l, height);
    // Clear the chosen flags with a kernel (no memset node if the caller captures).
    FillU32((uint32_t*)chosen, 0, (2ull * nleaves + 3) / 4, s);
    KPmaChoose<<<GridFor64(nleaves), 256, 0, s>>>(ins_cnt, S, nleaves, height, chosen, counters);
}
void LaunchPmaRebalance(const uint8_t* chosen, uint32_t nleaves, uint32_t height, uint32_t* wins,
                        unsigned long long* nwin, uint32_t* win_of, const uint32_t* gpos, const uint32_t* ins_leaf,
                        const uint64_t* ins_keys, const unsigned long long* n_ins_dev, uint64_t n_upper, PmaWin* meta,
                        uint64_t* slot, uint64_t* shadow, uint32_t* cnt, uint64_t* lo, unsigned long long* level_seen,
                        hipStream_t s) {
    FillU32(win_of, 0xFFFFFFFFu, nleaves, s);
    KPmaTop<<<GridFor64(2ull * nleaves), 256, 0, s>>>(chosen, nleaves, height, wins, nwin, win_of, level_seen);
    KPmaWinMeta<<<GridFor64(n_upper + 1), 256, 0, s>>>(wins, nwin, gpos, ins_leaf, n_ins_dev, meta);
    const uint64_t slots = (uint64_t)nleaves * kPmaLeaf;
    KPmaGather<<<GridFor64(slots), 256, 0, s>>>(slot, cnt, gpos, win_of, meta, shadow, nleaves);
    KPmaPlace<<<GridFor64(slots + n_upper), 256, 0, s>>>(shadow, ins_keys, ins_leaf, n_ins_dev, win_of, meta, slot,
                                                         nleaves);
    KPmaFinalize<<<std::min<uint32_t>(nleaves, 16384), 256, 0, s>>>(win_of, meta, slot, cnt, lo, nleaves);
}
void LaunchPmaDelete(const uint64_t* dkeys, const uint32_t* dleaf, uint64_t n, uint64_t* slot, uint32_t* cnt,
                     hipStream_t s) {
    if (n) KPmaDelete<<<GridFor64(n, 4, 16384), 256, 0, s>>>(dkeys, dleaf, n, slot, cnt);
}
void LaunchPmaFlat(const uint64_t* slot, const uint32_t* cnt, const uint32_t* gpos, uint32_t nleaves, uint64_t* flat,
                   hipStream_t s) {
    KPmaFlat<<<GridFor64((uint64_t)nleaves * kPmaLeaf), 256, 0, s>>>(slot, cnt, gpos, nleaves, flat);
}
void LaunchPmaRows(const uint64_t* flat, uint64_t n, uint32_t* pool, MRec* m_rec, uint64_t rows, hipStream_t s) {
    if (rows) KPmaClearRecs<<<GridFor64(rows), 256, 0, s>>>(m_rec, rows);
    if (n) KPmaRows<<<GridFor64(n), 256, 0, s>>>(flat, n, pool, m_rec);
}

}  // namespace cgd
