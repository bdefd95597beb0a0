// Index maintenance on gfx950: key decode, unique, segment discovery, and the
// per-list sorted merges that keep the main index M and build the update
// index U.  Replaces generateSizePtr (index/calig_helper_gamma.cu:223-236),
// mergeCSROntoGraph / removeCSROromGraph (kernels/indexing.cu:296-458) and
// the initial build of ConvertGlobalIndex (index/calig_helper_gamma.cu:238-360).
//
// Set semantics on insert (as index/cpu_index_mirror.cpp:155): values already
// present are dropped before the merge, so lists never hold duplicates.
//
// A touched list whose old length plus delta length is at most kShortWork
// (nearly all of them: lists hold a handful of entries) is merged by ONE lane,
// many lists per wavefront (KMergeInsertShort / KMergeRemoveShort):
//   insert: a forward two-pointer pass counts the delta values not yet
//   present, then one back-to-front pass writes the merged list (in place when
//   it fits its capacity: every store lands at or above the reads still to
//   come; else into a fresh pool block, the untouched head copied after);
//   remove: one forward compaction pass (stores land at or below the reads).
// Longer lists are queued (a ballot append) for one wavefront each:
//   1. dedup: each lane binary-searches its delta value in the old list; the
//      survivors B' are compacted with a ballot and their final positions
//      (rank in old list + rank in B') are recorded.
//   2. if the list keeps fitting its capacity, only the tail of the old list
//      from the first insertion point moves, back to front in 64-wide chunks
//      (every store lands at or above the loads still pending); otherwise the
//      merged list is written to a fresh pool block of CapFor(len) slots.
//   3. B' values are scattered to their recorded positions.
// Removal is an in-place ballot compaction starting at the first deleted value.
#include "engine.h"
#include "kernels.h"
#include "wave.h"

namespace cgd {

namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;

__global__ void KDecode(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n, uint32_t ushift,
                        uint32_t umask, uint32_t ndir, uint32_t vbits) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k = in[i];
        uint64_t e = k >> ushift, u = (k >> 32) & umask, v = (uint32_t)k;
        out[i] = ((u * ndir + e) << vbits) | v;
    }
}

__global__ void KMarkUnique(const uint64_t* __restrict__ keys, uint64_t n, uint32_t* __restrict__ flags) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void KCompactUnique(const uint64_t* __restrict__ keys, uint64_t n, const uint32_t* __restrict__ pos,
                               uint64_t* __restrict__ out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (i == 0 || keys[i] != keys[i - 1]) out[pos[i]] = keys[i];
}

__global__ void KMarkHeads(const uint64_t* __restrict__ keys, const unsigned long long* __restrict__ n_dev,
                           uint64_t n_upper, uint32_t vbits, uint32_t* __restrict__ flags) {
    const uint64_t n = *n_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_upper;
         i += (uint64_t)gridDim.x * blockDim.x)
        flags[i] = (i < n && (i == 0 || (keys[i] >> vbits) != (keys[i - 1] >> vbits))) ? 1u : 0u;
}

__global__ void KWriteSegs(const uint64_t* __restrict__ keys, const unsigned long long* __restrict__ n_dev,
                           uint32_t vbits, const uint32_t* __restrict__ pos, const unsigned long long* __restrict__ n_segs,
                           uint32_t* __restrict__ seg_row, uint64_t* __restrict__ seg_start) {
    const uint64_t n = *n_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == 0 || (keys[i] >> vbits) != (keys[i - 1] >> vbits)) {
            seg_row[pos[i]] = (uint32_t)(keys[i] >> vbits);
            seg_start[pos[i]] = i;
        }
        if (i == n - 1) seg_start[*n_segs] = n;
    }
}

__device__ __forceinline__ uint32_t LowerBoundKeys(const uint64_t* __restrict__ k, uint32_t n, uint32_t v,
                                                   uint64_t vmask) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t m = (lo + hi) >> 1;
        if ((uint32_t)(k[m] & vmask) < v) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// Lower bound of a wave-uniform v in sorted a[0, n): each level loads 64
// evenly spaced pivots (one per lane) and keeps the bucket that holds the
// answer, so a list of n entries costs ceil(log64 n) dependent round trips
// instead of log2 n.  Every lane must call it, with the same v.
__device__ __forceinline__ uint32_t WaveLowerBound(const uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
    const uint32_t lane = LaneId();
    uint32_t lo = 0;
    while (n > (uint32_t)kWave) {
        const uint32_t step = (n + kWave - 1) / kWave;
        const bool below = lane * step < n && a[lo + lane * step] < v;
        const uint32_t c = (uint32_t)__popcll(__ballot(below));
        if (c == 0) return lo;
        // the answer lies after the last pivot below v, up to the next pivot
        const uint32_t first = lo + (c - 1) * step + 1;
        const uint32_t end = c * step < n ? lo + c * step : lo + n;
        lo = first;
        n = end - first;
    }
    const bool below = lane < n && a[lo + lane] < v;
    return lo + (uint32_t)__popcll(__ballot(below));
}

// Lower bound of each lane's own v in sorted a[0, n), given the wave's 64
// pivots held = a[lane * step] (~0 past the end, step = ceil(n / 64)): the
// pivots below v name a bucket of step - 1 entries, searched alone.  The
// shuffle search counts at most 63 pivots; lane 63's is checked apart.
// Every lane must call it.
__device__ __forceinline__ uint32_t LowerBoundPiv(const uint32_t* __restrict__ a, uint32_t n, uint32_t step,
                                                  uint32_t held, uint32_t v) {
    const uint32_t c = RankInLanes(held, v) + (ReadLane(held, kWave - 1) < v ? 1u : 0u);
    if (c == 0) return 0;
    const uint32_t first = (c - 1) * step + 1;
    const uint32_t end = c * step < n ? c * step : n;
    return first + LowerBound(a + first, end - first, v);
}

// LowerBoundPiv over the values of sorted keys k[0, n) (value = key & vmask),
// given held = the value of k[lane * step] (~0 past the end).
__device__ __forceinline__ uint32_t LowerBoundKeysPiv(const uint64_t* __restrict__ k, uint32_t n, uint32_t step,
                                                      uint32_t held, uint32_t v, uint64_t vmask) {
    const uint32_t c = RankInLanes(held, v) + (ReadLane(held, kWave - 1) < v ? 1u : 0u);
    if (c == 0) return 0;
    const uint32_t first = (c - 1) * step + 1;
    const uint32_t end = c * step < n ? c * step : n;
    return first + LowerBoundKeys(k + first, end - first, v, vmask);
}

// The pivots of LowerBoundPiv for sorted a[0, n): a[lane * step].
__device__ __forceinline__ uint32_t Pivots(const uint32_t* __restrict__ a, uint32_t n, uint32_t step) {
    const uint32_t p = LaneId() * step;
    return p < n ? a[p] : ~0u;
}

constexpr uint32_t kShortWork = 48;
// Chunks of 64 entries a long-list merge loads per step (independent loads in
// flight per lane; a hub's list is thousands of chunks long).
constexpr uint32_t kMergeUnroll = 4;

// Appends segment s (lanes with push set) to the queue of long lists.
__device__ __forceinline__ void PushLong(bool push, uint64_t s, uint32_t* __restrict__ long_segs,
                                         unsigned long long* __restrict__ n_long) {
    const uint64_t m = __ballot(push);
    unsigned long long base = 0;
    if (m && LaneId() == 0) base = atomicAdd(n_long, (unsigned long long)__popcll(m));
    base = ReadLane64(base, 0);
    if (push) long_segs[base + RankBelow(m)] = (uint32_t)s;
}

__device__ __forceinline__ void FoldCounters(unsigned long long bytes, unsigned long long touched,
                                             Counters* __restrict__ cnt) {
    bytes = ReadLane64(WaveInclusiveScan(bytes), 63);
    touched = ReadLane64(WaveInclusiveScan(touched), 63);
    if (LaneId() == 0) {
        if (bytes) atomicAdd(&cnt->bytes, bytes);
        if (touched) atomicAdd(&cnt->rows_touched, touched);
    }
}

// One lane per short touched list: M[row] <- M[row] u B (B sorted, unique).
__global__ void __launch_bounds__(kThreads) KMergeInsertShort(Tables T, const uint64_t* __restrict__ keys,
                                                              uint32_t vbits, const uint32_t* __restrict__ seg_row,
                                                              const uint64_t* __restrict__ seg_start,
                                                              const unsigned long long* __restrict__ n_segs_dev,
                                                              uint64_t pool_cap, uint32_t* __restrict__ long_segs,
                                                              Counters* __restrict__ cnt) {
    const uint64_t n_segs = *n_segs_dev;
    const uint64_t vmask = (1ull << vbits) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bytes = 0, touched = 0;
    // The base is uniform over the block, so every lane reaches the ballot.
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n_segs; base += stride) {
        const uint64_t s = base + threadIdx.x;
        const bool valid = s < n_segs;
        uint32_t row = 0, nb = 0;
        uint64_t bs = 0;
        MRec rec{0, 0, 0};
        if (valid) {
            row = seg_row[s];
            bs = seg_start[s];
            nb = (uint32_t)(seg_start[s + 1] - bs);
            rec = T.m_rec[row];
        }
        const bool is_long = valid && rec.len + nb > kShortWork;
        PushLong(is_long, s, long_segs, &cnt->n_long);
        if (!valid || is_long) continue;
        const uint64_t* B = keys + bs;
        uint32_t* A = T.pool + rec.off;
        const uint32_t na = rec.len;
        // Forward pass: delta values not yet in the list.
        uint32_t nbp = 0;
        for (uint32_t j = 0, i = 0; j < nb; j++) {
            const uint32_t b = (uint32_t)(B[j] & vmask);
            while (i < na && A[i] < b) i++;
            nbp += (i < na && A[i] == b) ? 0u : 1u;
        }
        bytes += 8ull * nb + 4ull * na;
        if (nbp == 0) continue;
        const uint32_t newlen = na + nbp;
        uint32_t* C = A;
        if (newlen > rec.cap) {
            const uint32_t ncap = CapFor(newlen);
            const unsigned long long off = atomicAdd(&cnt->pool_top, (unsigned long long)ncap);
            if (off + ncap > pool_cap) {
                atomicOr(&cnt->flags, kFlagPool);
                continue;
            }
            C = T.pool + off;
            rec.off = off;
            rec.cap = ncap;
        }
        // Back-to-front merge; w - i = delta values still to place.
        uint32_t i = na, w = newlen;
        for (uint32_t j = nb; j > 0; j--) {
            const uint32_t b = (uint32_t)(B[j - 1] & vmask);
            while (i > 0) {
                const uint32_t a = A[i - 1];
                if (a <= b) break;
                C[--w] = a;
                i--;
            }
            if (i > 0 && A[i - 1] == b) continue;
            C[--w] = b;
        }
        bytes += 8ull * (na - i) + 4ull * nbp;
        if (C != A)
            for (uint32_t k = 0; k < i; k++) C[k] = A[k];
        if (C != A) bytes += 8ull * i;
        rec.len = newlen;
        T.m_rec[row] = rec;
        touched++;
    }
    FoldCounters(bytes, touched, cnt);
}

// One lane per short touched list: M[row] <- M[row] \ D (D sorted).
__global__ void __launch_bounds__(kThreads) KMergeRemoveShort(Tables T, const uint64_t* __restrict__ keys,
                                                              uint32_t vbits, const uint32_t* __restrict__ seg_row,
                                                              const uint64_t* __restrict__ seg_start,
                                                              const unsigned long long* __restrict__ n_segs_dev,
                                                              uint32_t* __restrict__ long_segs,
                                                              Counters* __restrict__ cnt) {
    const uint64_t n_segs = *n_segs_dev;
    const uint64_t vmask = (1ull << vbits) - 1ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bytes = 0, touched = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n_segs; base += stride) {
        const uint64_t s = base + threadIdx.x;
        const bool valid = s < n_segs;
        uint32_t row = 0, nd = 0, na = 0;
        uint64_t bs = 0, off = 0;
        if (valid) {
            row = seg_row[s];
            bs = seg_start[s];
            nd = (uint32_t)(seg_start[s + 1] - bs);
            const MRec rec = T.m_rec[row];
            na = rec.len;
            off = rec.off;
        }
        const bool is_long = valid && na + nd > kShortWork;
        PushLong(is_long, s, long_segs, &cnt->n_long);
        if (!valid || is_long) continue;
        bytes += 8ull * nd;
        if (na == 0) continue;
        const uint64_t* D = keys + bs;
        uint32_t* A = T.pool + off;
        uint32_t w = 0;
        for (uint32_t i = 0, j = 0; i < na; i++) {
            const uint32_t a = A[i];
            while (j < nd && (uint32_t)(D[j] & vmask) < a) j++;
            if (j < nd && (uint32_t)(D[j] & vmask) == a) continue;
            if (w != i) A[w] = a;
            w++;
        }
        bytes += 4ull * na + 4ull * w;
        if (w != na) {
            touched++;
            T.m_rec[row].len = w;
        }
    }
    FoldCounters(bytes, touched, cnt);
}

// One wavefront per long touched list: M[row] <- M[row] u B (B sorted, unique).
__global__ void __launch_bounds__(kThreads) KMergeInsert(Tables T, const uint64_t* __restrict__ keys, uint32_t vbits,
                                                         const uint32_t* __restrict__ seg_row,
                                                         const uint64_t* __restrict__ seg_start,
                                                         const uint32_t* __restrict__ long_segs,
                                                         const unsigned long long* __restrict__ n_long_dev,
                                                         uint32_t* __restrict__ sc_val, uint32_t* __restrict__ sc_pos,
                                                         uint64_t pool_cap, Counters* __restrict__ cnt) {
    const uint32_t lane = LaneId();
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    const uint64_t n_long = *n_long_dev;
    const uint64_t vmask = (1ull << vbits) - 1ull;
    unsigned long long bytes = 0, touched = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; t < n_long; t += nw) {
        const uint64_t s = long_segs[t];
        const uint32_t row = seg_row[s];
        const uint64_t bs = seg_start[s];
        const uint32_t nb = (uint32_t)(seg_start[s + 1] - bs);
        MRec rec = T.m_rec[row];
        uint32_t* A = T.pool + rec.off;
        const uint32_t na = rec.len;
        // 1. B' = B \ A with final positions (lists of more than 128
        // entries searched from 64 pivots).
        const bool a_piv = na > 2u * kWave;
        const uint32_t astep = (na + kWave - 1) / kWave;
        const uint32_t apiv = a_piv ? Pivots(A, na, astep) : 0u;
        uint32_t nbp = 0;
        for (uint32_t k = 0; k < nb; k += kWave) {
            const uint32_t j = k + lane;
            const bool valid = j < nb;
            const uint32_t b = valid ? (uint32_t)(keys[bs + j] & vmask) : 0u;
            const uint32_t p = a_piv ? LowerBoundPiv(A, na, astep, apiv, b) : (valid ? LowerBound(A, na, b) : 0u);
            const bool keep = valid && !(p < na && A[p] == b);
            const uint64_t m = __ballot(keep);
            const uint32_t r = RankBelow(m);
            if (keep) {
                sc_val[bs + nbp + r] = b;
                sc_pos[bs + nbp + r] = p + nbp + r;
            }
            nbp += (uint32_t)__popcll(m);
        }
        bytes += 8ull * nb;
        if (nbp == 0) continue;
        WaveSync();
        const uint32_t newlen = na + nbp;
        const uint32_t* Bv = sc_val + bs;
        const uint32_t* Bp = sc_pos + bs;
        // Fewer than 64 new values: held one per lane, so a moved entry's
        // shift (new values below it) is found by shuffles; more: 64 pivots
        // held one per lane narrow each search to one bucket.
        const bool in_lanes = nbp < (uint32_t)kWave;
        const uint32_t bstep = (nbp + kWave - 1) / kWave;
        const uint32_t held = Pivots(Bv, nbp, bstep);
        uint32_t* C = A;
        if (newlen > rec.cap) {
            const uint32_t ncap = CapFor(newlen);
            unsigned long long off = 0;
            if (lane == 0) off = atomicAdd(&cnt->pool_top, (unsigned long long)ncap);
            off = ReadLane64(off, 0);
            if (off + ncap > pool_cap) {
                if (lane == 0) atomicOr(&cnt->flags, kFlagPool);
                continue;
            }
            C = T.pool + off;
            for (uint32_t k = 0; k < na; k += kWave * kMergeUnroll) {
                uint32_t a[kMergeUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kMergeUnroll; u++) {
                    const uint32_t i = k + u * kWave + lane;
                    a[u] = i < na ? A[i] : 0u;
                }
#pragma unroll
                for (uint32_t u = 0; u < kMergeUnroll; u++) {
                    const uint32_t i = k + u * kWave + lane;
                    const uint32_t r = in_lanes ? RankInLanes(held, a[u]) : LowerBoundPiv(Bv, nbp, bstep, held, a[u]);
                    if (i < na) C[i + r] = a[u];
                }
            }
            bytes += 4ull * na + 4ull * na;
            if (lane == 0) {
                rec.off = off;
                rec.cap = ncap;
            }
        } else {
            // In place: only A[i0..na) moves, back to front, kMergeUnroll
            // chunks of 64 per step, all loaded before any is stored.  An
            // entry moves up by the new values below it, which never exceeds
            // the shift of the entries above it, so a store lands at or above
            // its own source (read this step) and below every entry moved in
            // earlier steps.
            const uint32_t i0 = Bp[0];
            for (uint32_t hi = na; hi > i0;) {
                const uint32_t lo = hi > i0 + kWave * kMergeUnroll ? hi - kWave * kMergeUnroll : i0;
                uint32_t a[kMergeUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kMergeUnroll; u++) {
                    const uint32_t d = u * kWave + lane;  // distance below hi - 1
                    a[u] = d < hi - lo ? A[hi - 1 - d] : 0u;
                }
                uint32_t r[kMergeUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kMergeUnroll; u++)
                    r[u] = in_lanes ? RankInLanes(held, a[u]) : LowerBoundPiv(Bv, nbp, bstep, held, a[u]);
                WaveSync();
#pragma unroll
                for (uint32_t u = 0; u < kMergeUnroll; u++) {
                    const uint32_t d = u * kWave + lane;
                    if (d < hi - lo) A[hi - 1 - d + r[u]] = a[u];
                }
                WaveSync();
                hi = lo;
            }
            bytes += 8ull * (na - i0);
        }
        for (uint32_t k = 0; k < nbp; k += kWave) {
            const uint32_t j = k + lane;
            if (j < nbp) C[Bp[j]] = Bv[j];
        }
        bytes += 4ull * nbp;
        touched++;
        if (lane == 0) {
            rec.len = newlen;
            T.m_rec[row] = rec;
        }
    }
    if (lane == 0) {
        atomicAdd(&cnt->bytes, bytes);
        atomicAdd(&cnt->rows_touched, touched);
    }
}

// One wavefront per long touched list: M[row] <- M[row] \ D (D sorted).
__global__ void __launch_bounds__(kThreads) KMergeRemove(Tables T, const uint64_t* __restrict__ keys, uint32_t vbits,
                                                         const uint32_t* __restrict__ seg_row,
                                                         const uint64_t* __restrict__ seg_start,
                                                         const uint32_t* __restrict__ long_segs,
                                                         const unsigned long long* __restrict__ n_long_dev,
                                                         Counters* __restrict__ cnt) {
    const uint32_t lane = LaneId();
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    const uint64_t n_long = *n_long_dev;
    const uint64_t vmask = (1ull << vbits) - 1ull;
    unsigned long long bytes = 0, touched = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; t < n_long; t += nw) {
        const uint64_t s = long_segs[t];
        const uint32_t row = seg_row[s];
        const uint64_t bs = seg_start[s];
        const uint32_t nd = (uint32_t)(seg_start[s + 1] - bs);
        const uint64_t* D = keys + bs;
        const uint32_t na = T.m_rec[row].len;
        uint32_t* A = T.pool + T.m_rec[row].off;
        bytes += 8ull * nd;
        if (na == 0) continue;
        const uint32_t i0 = WaveLowerBound(A, na, (uint32_t)(D[0] & vmask));
        uint32_t w = i0;
        // Fewer than 64 removed values: held one per lane (see KMergeInsert);
        // more: 64 pivots held one per lane narrow each search to a bucket.
        const bool in_lanes = nd < (uint32_t)kWave;
        const uint32_t dstep = (nd + kWave - 1) / kWave;
        const uint32_t held = lane * dstep < nd ? (uint32_t)(D[lane * dstep] & vmask) : ~0u;
        // kMergeUnroll chunks of 64 are loaded before any of them is stored:
        // the kept entries move down to w <= k, so a store never reaches an
        // entry of a later step, and every entry of this step is in registers.
        for (uint32_t k = i0; k < na; k += kWave * kMergeUnroll) {
            uint32_t a[kMergeUnroll];
            bool keep[kMergeUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kMergeUnroll; u++) {
                const uint32_t i = k + u * kWave + lane;
                keep[u] = i < na;
                a[u] = keep[u] ? A[i] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < kMergeUnroll; u++) {
                bool del = false;
                if (in_lanes) {
                    // every lane shuffles (a lane left out would not supply its value)
                    const uint32_t p = RankInLanes(held, a[u]);
                    const uint32_t at = __shfl(held, (int)(p < (uint32_t)kWave ? p : kWave - 1), kWave);
                    del = keep[u] && p < nd && at == a[u];
                } else {
                    const uint32_t p = LowerBoundKeysPiv(D, nd, dstep, held, a[u], vmask);
                    del = keep[u] && p < nd && (uint32_t)(D[p] & vmask) == a[u];
                }
                keep[u] = keep[u] && !del;
            }
            WaveSync();
#pragma unroll
            for (uint32_t u = 0; u < kMergeUnroll; u++) {
                const uint64_t m = __ballot(keep[u]);
                if (keep[u]) A[w + RankBelow(m)] = a[u];
                w += (uint32_t)__popcll(m);
            }
            WaveSync();
        }
        bytes += 4ull * (na - i0) + 4ull * (w - i0);
        if (w != na) touched++;
        if (lane == 0) T.m_rec[row].len = w;
    }
    if (lane == 0) {
        atomicAdd(&cnt->bytes, bytes);
        atomicAdd(&cnt->rows_touched, touched);
    }
}

__global__ void KClearU(URec* __restrict__ u_rec, const uint32_t* __restrict__ u_rows,
                        const unsigned long long* __restrict__ n_dev) {
    const uint64_t n = *n_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        u_rec[u_rows[i]] = URec{0, 0};
}

__global__ void KBuildUValues(const uint64_t* __restrict__ keys, uint64_t n, uint32_t vbits, uint32_t* __restrict__ vals) {
    const uint64_t vmask = (1ull << vbits) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        vals[i] = (uint32_t)(keys[i] & vmask);
}

__global__ void KBuildURows(URec* __restrict__ u_rec, const uint32_t* __restrict__ seg_row,
                            const uint64_t* __restrict__ seg_start, const unsigned long long* __restrict__ n_segs_dev,
                            uint32_t* __restrict__ u_rows, unsigned long long* __restrict__ u_rows_count) {
    const uint64_t n = *n_segs_dev;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t row = seg_row[s];
        u_rec[row] = URec{(uint32_t)seg_start[s], (uint32_t)(seg_start[s + 1] - seg_start[s])};
        u_rows[s] = row;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *u_rows_count = n;
}

// Initial build: per segment capacity (in the u64 scan buffer), then layout.
__global__ void KInitCaps(const uint64_t* __restrict__ seg_start, const unsigned long long* __restrict__ n_segs_dev,
                          uint64_t n_upper, uint64_t* __restrict__ caps) {
    const uint64_t n = *n_segs_dev;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n_upper; s += (uint64_t)gridDim.x * blockDim.x)
        caps[s] = s < n ? CapFor((uint32_t)(seg_start[s + 1] - seg_start[s])) : 0;
}

__global__ void KInitRecs(Tables T, const uint32_t* __restrict__ seg_row, const uint64_t* __restrict__ seg_start,
                          const unsigned long long* __restrict__ n_segs_dev, const uint64_t* __restrict__ offs,
                          Counters* __restrict__ cnt) {
    const uint64_t n = *n_segs_dev;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t row = seg_row[s];
        const uint32_t len = (uint32_t)(seg_start[s + 1] - seg_start[s]);
        T.m_rec[row] = MRec{offs[s], len, CapFor(len)};
        const uint32_t e = row % T.ndir;
        atomicAdd(&cnt->card[e], (unsigned long long)len);
        atomicAdd(&cnt->nodes[e], 1ull);
    }
}

__global__ void KInitValues(Tables T, const uint64_t* __restrict__ keys, const unsigned long long* __restrict__ n_dev,
                            uint32_t vbits, const uint32_t* __restrict__ head_pos, const uint64_t* __restrict__ seg_start,
                            const uint64_t* __restrict__ offs) {
    const uint64_t n = *n_dev;
    const uint64_t vmask = (1ull << vbits) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const bool head = i == 0 || (keys[i] >> vbits) != (keys[i - 1] >> vbits);
        const uint64_t seg = (uint64_t)head_pos[i] + (head ? 1 : 0) - 1;
        T.pool[offs[seg] + (i - seg_start[seg])] = (uint32_t)(keys[i] & vmask);
    }
}

// Export helpers: M/U as (row, v) pairs in boundary key form, row order.
__global__ void KExportM(Tables T, const uint64_t* __restrict__ row_off, uint64_t nrows, uint32_t ushift,
                         uint64_t* __restrict__ out) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (uint64_t)gridDim.x * blockDim.x) {
        const MRec rec = T.m_rec[r];
        const uint64_t e = r % T.ndir, u = r / T.ndir;
        for (uint32_t j = 0; j < rec.len; j++)
            out[row_off[r] + j] = (e << ushift) | (u << 32) | T.pool[rec.off + j];
    }
}

__global__ void KRowLens(Tables T, uint64_t nrows, int which, uint64_t* __restrict__ lens) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (uint64_t)gridDim.x * blockDim.x)
        lens[r] = which ? T.u_rec[r].len : T.m_rec[r].len;
}

__global__ void KExportU(Tables T, const uint64_t* __restrict__ row_off, uint64_t nrows, uint32_t ushift,
                         uint64_t* __restrict__ out) {
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += (uint64_t)gridDim.x * blockDim.x) {
        const URec rec = T.u_rec[r];
        const uint64_t e = r % T.ndir, u = r / T.ndir;
        for (uint32_t j = 0; j < rec.len; j++)
            out[row_off[r] + j] = (e << ushift) | (u << 32) | T.u_vals[rec.off + j];
    }
}

unsigned GridFor(uint64_t n, unsigned per_block = 256, unsigned cap = 8192) {
    uint64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    return (unsigned)std::min<uint64_t>(g, cap);
}

}  // namespace

void LaunchDecode(const uint64_t* in, uint64_t* out, uint64_t n, uint32_t ushift, uint32_t umask, uint32_t ndir,
                  uint32_t vbits, hipStream_t s) {
    if (!n) return;
    KDecode<<<GridFor(n), 256, 0, s>>>(in, out, n, ushift, umask, ndir, vbits);
}

void LaunchUnique(const uint64_t* in, uint64_t* out, uint64_t n, uint32_t* flags, void* scan_scratch,
                  unsigned long long* n_out_dev, hipStream_t s) {
    if (!n) {
        FillU64((uint64_t*)n_out_dev, 0, 1, s);
        return;
    }
    KMarkUnique<<<GridFor(n), 256, 0, s>>>(in, n, flags);
    ScanU32(flags, flags, n, (uint64_t*)n_out_dev, scan_scratch, s);
    KCompactUnique<<<GridFor(n), 256, 0, s>>>(in, n, flags, out);
}

void LaunchSegments(const uint64_t* keys, const unsigned long long* n_dev, uint64_t n_upper, uint32_t vbits,
                    uint32_t* flags, uint32_t* seg_row, uint64_t* seg_start, unsigned long long* n_segs_dev,
                    void* scan_scratch, hipStream_t s) {
    if (!n_upper) {
        FillU64((uint64_t*)n_segs_dev, 0, 1, s);
        FillU64(seg_start, 0, 1, s);
        return;
    }
    KMarkHeads<<<GridFor(n_upper), 256, 0, s>>>(keys, n_dev, n_upper, vbits, flags);
    ScanU32(flags, flags, n_upper, (uint64_t*)n_segs_dev, scan_scratch, s);
    FillU64(seg_start, 0, 1, s);  // sentinel when the key set is empty
    KWriteSegs<<<GridFor(n_upper), 256, 0, s>>>(keys, n_dev, vbits, flags, n_segs_dev, seg_row, seg_start);
}

// Both merges expect cnt->n_long zeroed; long_segs holds n_upper entries.
void LaunchMergeInsert(const Tables& T, const uint64_t* keys, uint32_t vbits, const uint32_t* seg_row,
                       const uint64_t* seg_start, const unsigned long long* n_segs, uint64_t n_upper, uint32_t* sc_val,
                       uint32_t* sc_pos, uint32_t* long_segs, uint64_t pool_cap, Counters* cnt, hipStream_t s) {
    if (!n_upper) return;
    KMergeInsertShort<<<GridFor(n_upper, kThreads, 8192), kThreads, 0, s>>>(T, keys, vbits, seg_row, seg_start,
                                                                           n_segs, pool_cap, long_segs, cnt);
    KMergeInsert<<<GridFor(n_upper, kWavesPerBlock, 2048), kThreads, 0, s>>>(
        T, keys, vbits, seg_row, seg_start, long_segs, &cnt->n_long, sc_val, sc_pos, pool_cap, cnt);
}

void LaunchMergeRemove(const Tables& T, const uint64_t* keys, uint32_t vbits, const uint32_t* seg_row,
                       const uint64_t* seg_start, const unsigned long long* n_segs, uint64_t n_upper,
                       uint32_t* long_segs, Counters* cnt, hipStream_t s) {
    if (!n_upper) return;
    KMergeRemoveShort<<<GridFor(n_upper, kThreads, 8192), kThreads, 0, s>>>(T, keys, vbits, seg_row, seg_start,
                                                                           n_segs, long_segs, cnt);
    KMergeRemove<<<GridFor(n_upper, kWavesPerBlock, 2048), kThreads, 0, s>>>(T, keys, vbits, seg_row, seg_start,
                                                                            long_segs, &cnt->n_long, cnt);
}

void LaunchClearU(URec* u_rec, const uint32_t* u_rows, const unsigned long long* n_dev, uint64_t n_upper,
                  hipStream_t s) {
    if (!n_upper) return;
    KClearU<<<GridFor(n_upper), 256, 0, s>>>(u_rec, u_rows, n_dev);
}

void LaunchBuildU(URec* u_rec, const uint64_t* keys, uint64_t n, uint32_t vbits, uint32_t* vals,
                  const uint32_t* seg_row, const uint64_t* seg_start, const unsigned long long* n_segs, uint32_t* u_rows,
                  unsigned long long* u_rows_count, hipStream_t s) {
    if (n) KBuildUValues<<<GridFor(n), 256, 0, s>>>(keys, n, vbits, vals);
    KBuildURows<<<GridFor(n ? n : 1), 256, 0, s>>>(u_rec, seg_row, seg_start, n_segs, u_rows, u_rows_count);
}

void LaunchInitCaps(const uint64_t* seg_start, const unsigned long long* n_segs, uint64_t n_upper, uint64_t* caps,
                    hipStream_t s) {
    if (!n_upper) return;
    KInitCaps<<<GridFor(n_upper), 256, 0, s>>>(seg_start, n_segs, n_upper, caps);
}

void LaunchInitLayout(const Tables& T, const uint64_t* keys, const unsigned long long* n_dev, uint64_t n_upper,
                      uint32_t vbits, const uint32_t* head_pos, const uint32_t* seg_row, const uint64_t* seg_start,
                      const unsigned long long* n_segs, const uint64_t* offs, Counters* cnt, hipStream_t s) {
    if (!n_upper) return;
    KInitRecs<<<GridFor(n_upper), 256, 0, s>>>(T, seg_row, seg_start, n_segs, offs, cnt);
    KInitValues<<<GridFor(n_upper), 256, 0, s>>>(T, keys, n_dev, vbits, head_pos, seg_start, offs);
}

void LaunchRowLens(const Tables& T, uint64_t nrows, int which, uint64_t* lens, hipStream_t s) {
    if (!nrows) return;
    KRowLens<<<GridFor(nrows), 256, 0, s>>>(T, nrows, which, lens);
}

void LaunchExport(const Tables& T, uint64_t nrows, int which, const uint64_t* row_off, uint32_t ushift, uint64_t* out,
                  hipStream_t s) {
    if (!nrows) return;
    if (which)
        KExportU<<<GridFor(nrows), 256, 0, s>>>(T, row_off, nrows, ushift, out);
    else
        KExportM<<<GridFor(nrows), 256, 0, s>>>(T, row_off, nrows, ushift, out);
}

// Adjacent equal keys of a sorted key array into cnt->u_dups (0 means the
// update lists are sets).
__global__ void KCountDups(const uint64_t* __restrict__ k, uint64_t n, Counters* __restrict__ cnt) {
    unsigned long long mine = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        mine += k[i] == k[i - 1];
    const unsigned long long tot = ReadLane64(WaveInclusiveScan(mine), 63);
    if (LaneId() == 0 && tot) atomicAdd(&cnt->u_dups, tot);
}

void LaunchCountDups(const uint64_t* k, uint64_t n, Counters* cnt, hipStream_t s) {
    if (n < 2) return;
    const unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 2048);
    KCountDups<<<g, 256, 0, s>>>(k, n, cnt);
}

}  // namespace cgd
