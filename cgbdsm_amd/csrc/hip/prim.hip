// Device-wide primitives for gfx950: exclusive scan (reduce-then-scan) and a
// stable LSD radix sort of u64 keys with 8-bit digits.  These replace the
// thrust::sort / unique and CUB scan calls of the reference
// (index/calig_helper_gamma.cu:304, 491-492, 573; graph/graph_gpu.cu:1395-1400).
#include "engine.h"
#include "wave.h"

namespace cgd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // 2048

template <class T>
__device__ T BlockExclusiveScan(T v, T* total) {
    __shared__ T wsum[kScanThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    T inc = WaveInclusiveScan(v);
    if (lane == kWave - 1) wsum[w] = inc;
    __syncthreads();
    T pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / kWave; i++) {
        if (i < w) pre += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - v;
}

template <class T>
__global__ void __launch_bounds__(kScanThreads) KScanReduce(const T* in, uint64_t n, T* tile_sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    T s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        uint64_t idx = base + (uint64_t)threadIdx.x * kScanItems + i;
        if (idx < n) s += in[idx];
    }
    T tot;
    BlockExclusiveScan(s, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Single block scans the tile sums in place and writes the grand total.
template <class T>
__global__ void __launch_bounds__(kScanThreads) KScanTiles(T* tile_sums, uint64_t nt, uint64_t* total) {
    T carry = 0;
    for (uint64_t base = 0; base < nt; base += kScanThreads) {
        uint64_t idx = base + threadIdx.x;
        T v = idx < nt ? tile_sums[idx] : 0;
        T tot;
        T ex = BlockExclusiveScan(v, &tot);
        if (idx < nt) tile_sums[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = (uint64_t)carry;
}

template <class T>
__global__ void __launch_bounds__(kScanThreads) KScanApply(const T* in, T* out, uint64_t n, const T* tile_sums) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    T v[kScanItems];
    T s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        uint64_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0;
        s += v[i];
    }
    T tot;
    T ex = BlockExclusiveScan(s, &tot) + tile_sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        uint64_t idx = base + i;
        if (idx < n) out[idx] = ex;
        ex += v[i];
    }
}

// Whole scan in one block, tile after tile with a running carry (in place
// allowed: a tile is read into registers before it is written).  For short
// inputs one launch instead of three: the per-batch index kernels run dozens
// of such scans per query.
template <class T>
__global__ void __launch_bounds__(kScanThreads) KScanSingle(const T* in, T* out, uint64_t n, uint64_t* total) {
    T carry = 0;
    for (uint64_t base = 0; base < n; base += kScanTile) {
        const uint64_t b0 = base + (uint64_t)threadIdx.x * kScanItems;
        T v[kScanItems];
        T sum = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            v[i] = b0 + i < n ? in[b0 + i] : 0;
            sum += v[i];
        }
        T tot;
        T ex = BlockExclusiveScan(sum, &tot) + carry;
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            if (b0 + i < n) out[b0 + i] = ex;
            ex += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = (uint64_t)carry;
}

constexpr uint64_t kScanSingleMax = 4 * kScanTile;  // 8192 elements (the block walks its tiles in turn)

template <class T>
void ScanImpl(const T* in, T* out, uint64_t n, uint64_t* total_dev, void* scratch, hipStream_t s) {
    if (n == 0) {
        if (total_dev) FillU64(total_dev, 0, 1, s);
        return;
    }
    if (n <= kScanSingleMax) {
        KScanSingle<T><<<1, kScanThreads, 0, s>>>(in, out, n, total_dev);
        CGD_HIP(hipGetLastError());
        return;
    }
    const uint64_t nt = (n + kScanTile - 1) / kScanTile;
    T* tiles = (T*)scratch;
    KScanReduce<T><<<(unsigned)nt, kScanThreads, 0, s>>>(in, n, tiles);
    KScanTiles<T><<<1, kScanThreads, 0, s>>>(tiles, nt, total_dev);
    KScanApply<T><<<(unsigned)nt, kScanThreads, 0, s>>>(in, out, n, tiles);
    CGD_HIP(hipGetLastError());
}

__global__ void KFill64(uint64_t* p, uint64_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void KSetWords(uint64_t* base, WordSet ws) {
    if (threadIdx.x < ws.n) base[ws.off[threadIdx.x]] = ws.val[threadIdx.x];
}
__global__ void KFill32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------- radix sort
constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;  // 4096 keys per block

__global__ void __launch_bounds__(kRsThreads) KRsHist(const uint64_t* keys, uint64_t n, uint32_t shift,
                                                       uint32_t* hist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
#pragma unroll 4
    for (int i = 0; i < kRsItems; i++) {
        uint64_t idx = base + (uint64_t)i * kRsThreads + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: within a block, keys are ranked round by round (256 keys
// per round, in index order); inside a wave the rank among equal digits comes
// from 8 ballots, across waves from per-wave digit counts in LDS.
__global__ void __launch_bounds__(kRsThreads) KRsScatter(const uint64_t* in, uint64_t* out, uint64_t n,
                                                          uint32_t shift, const uint32_t* offs, uint32_t nblocks) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kRsThreads / kWave][256];
    const int tid = threadIdx.x, w = tid / kWave;
    const uint32_t lane = tid & (kWave - 1);
    run[tid] = offs[(uint64_t)tid * nblocks + blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; r++) {
#pragma unroll
        for (int ww = 0; ww < kRsThreads / kWave; ww++) wcnt[ww][tid] = 0;
        __syncthreads();
        const uint64_t idx = base + (uint64_t)r * kRsThreads + tid;
        const bool valid = idx < n;
        const uint64_t key = valid ? in[idx] : 0;
        const uint32_t d = (uint32_t)(key >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            uint64_t bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = __popcll(m & LaneMaskLt(lane));
        if (valid && rank == 0) wcnt[w][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t pre = run[d];
            for (int ww = 0; ww < w; ww++) pre += wcnt[ww][d];
            out[pre + rank] = key;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int ww = 0; ww < kRsThreads / kWave; ww++) add += wcnt[ww][tid];
        run[tid] += add;
        __syncthreads();
    }
}

// Whole sort of at most one tile (kRsTile keys) in one block: the keys stay
// in LDS across the digit passes; each pass takes the digit histogram, its
// exclusive scan, then the same stable round-by-round scatter as KRsScatter
// (the block's keys held in registers while LDS is rewritten).  One launch
// instead of three per digit for the small key sets of most batches.
__global__ void __launch_bounds__(kRsThreads) KRsSortTile(uint64_t* keys, uint32_t n, uint32_t bits) {
    __shared__ uint64_t buf[kRsTile];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kRsThreads / kWave][256];
    const int tid = threadIdx.x, w = tid / kWave;
    const uint32_t lane = tid & (kWave - 1);
    for (uint32_t i = tid; i < n; i += kRsThreads) buf[i] = keys[i];
    __syncthreads();
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        run[tid] = 0;
        __syncthreads();
        for (uint32_t i = tid; i < n; i += kRsThreads) atomicAdd(&run[(uint32_t)(buf[i] >> shift) & 255u], 1u);
        __syncthreads();
        // exclusive scan of the 256 digit counts (Hillis-Steele in LDS)
        uint32_t v = run[tid];
        const uint32_t own = v;
        for (int o = 1; o < 256; o <<= 1) {
            __syncthreads();
            const uint32_t add = tid >= o ? run[tid - o] : 0u;
            __syncthreads();
            v += add;
            run[tid] = v;
        }
        __syncthreads();
        run[tid] = v - own;
        uint64_t k[kRsItems];
#pragma unroll
        for (int r = 0; r < kRsItems; r++) {
            const uint32_t idx = (uint32_t)r * kRsThreads + tid;
            k[r] = idx < n ? buf[idx] : 0;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRsItems; r++) {
#pragma unroll
            for (int ww = 0; ww < kRsThreads / kWave; ww++) wcnt[ww][tid] = 0;
            __syncthreads();
            const uint32_t idx = (uint32_t)r * kRsThreads + tid;
            const bool valid = idx < n;
            const uint64_t key = k[r];
            const uint32_t d = (uint32_t)(key >> shift) & 255u;
            uint64_t m = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                uint64_t bb = __ballot((d >> b) & 1u);
                m &= ((d >> b) & 1u) ? bb : ~bb;
            }
            const uint32_t rank = __popcll(m & LaneMaskLt(lane));
            if (valid && rank == 0) wcnt[w][d] = __popcll(m);
            __syncthreads();
            if (valid) {
                uint32_t pre = run[d];
                for (int ww = 0; ww < w; ww++) pre += wcnt[ww][d];
                buf[pre + rank] = key;
            }
            __syncthreads();
            uint32_t add = 0;
#pragma unroll
            for (int ww = 0; ww < kRsThreads / kWave; ww++) add += wcnt[ww][tid];
            run[tid] += add;
            __syncthreads();
        }
    }
    for (uint32_t i = tid; i < n; i += kRsThreads) keys[i] = buf[i];
}

}  // namespace

size_t ScanScratchBytes(uint64_t n) { return ((n + kScanTile - 1) / kScanTile + 1) * sizeof(uint64_t); }

void ScanU32(const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* total_dev, void* scratch, hipStream_t s) {
    ScanImpl<uint32_t>(in, out, n, total_dev, scratch, s);
}
void ScanU64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* total_dev, void* scratch, hipStream_t s) {
    ScanImpl<uint64_t>(in, out, n, total_dev, scratch, s);
}

void FillU64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
    if (!n) return;
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    KFill64<<<g, 256, 0, s>>>(p, v, n);
}
void SetWords(uint64_t* base, const WordSet& ws, hipStream_t s) {
    if (ws.n) KSetWords<<<1, 64, 0, s>>>(base, ws);
}
void FillU32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (!n) return;
    unsigned g = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    KFill32<<<g, 256, 0, s>>>(p, v, n);
}

size_t RadixHistWords(uint64_t n) {
    uint64_t nb = (n + kRsTile - 1) / kRsTile;
    uint64_t hn = (nb ? nb : 1) * 256;
    // histogram, one pad word, then the scan scratch of the histogram
    return (size_t)(hn + 1 + ScanScratchBytes(hn) / sizeof(uint32_t) + 2);
}

void RadixSortU64(uint64_t* keys, uint64_t* alt, uint64_t n, uint32_t bits, uint32_t* hist, hipStream_t s) {
    if (n < 2) return;
    if (n <= (uint64_t)kRsTile) {
        KRsSortTile<<<1, kRsThreads, 0, s>>>(keys, (uint32_t)n, bits);
        CGD_HIP(hipGetLastError());
        return;
    }
    const uint32_t nb = (uint32_t)((n + kRsTile - 1) / kRsTile);
    const uint64_t hn = (uint64_t)nb * 256;
    // The scan scratch lives right after the histogram.
    uint32_t* scan_scratch = hist + hn + 1;
    uint64_t* src = keys;
    uint64_t* dst = alt;
    for (uint32_t shift = 0; shift < bits; shift += 8) {
        KRsHist<<<nb, kRsThreads, 0, s>>>(src, n, shift, hist, nb);
        ScanU32(hist, hist, hn, nullptr, scan_scratch, s);
        KRsScatter<<<nb, kRsThreads, 0, s>>>(src, dst, n, shift, hist, nb);
        std::swap(src, dst);
    }
    if (src != keys) CGD_HIP(hipMemcpyAsync(keys, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
    CGD_HIP(hipGetLastError());
}

}  // namespace cgd
