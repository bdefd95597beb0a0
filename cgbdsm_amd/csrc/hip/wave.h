// wave64 helpers for gfx950 (CDNA4): lane masks, DPP scans and reductions,
// lane reads, and an intra-wave LDS visibility fence.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cgd {

__device__ __forceinline__ uint32_t LaneId() { return __lane_id(); }

__device__ __forceinline__ uint64_t LaneMaskLt(uint32_t lane) { return (1ull << lane) - 1ull; }

// Number of set bits of `mask` below the calling lane (v_mbcnt).
__device__ __forceinline__ uint32_t RankBelow(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Orders LDS and global traffic between lanes of one wavefront (lanes run in
// lockstep, but the memory model needs the fence to publish stores).
__device__ __forceinline__ void WaveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// DPP lane moves (GFX9 wave64 controls, VALU only: no LDS-pipe round trip
// as a shuffle takes).  kDppShr<n> moves lane i - n of the same row of 16 to
// lane i; kDppBcast15 (rows 1 and 3 enabled) moves lane 15 of the previous
// row, kDppBcast31 (rows 2 and 3) moves lane 31.  A lane whose source lies
// outside its row, or whose row is not enabled, reads `fill`.  Every lane of
// the wavefront must be active.
constexpr int kDppShr1 = 0x111, kDppShr2 = 0x112, kDppShr4 = 0x114, kDppShr8 = 0x118;
constexpr int kDppBcast15 = 0x142, kDppBcast31 = 0x143;
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ uint32_t DppMove(uint32_t v, uint32_t fill = 0u) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)fill, (int)v, Ctrl, RowMask, 0xf, false);
}
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ uint64_t DppMove64(uint64_t v) {
    return ((uint64_t)DppMove<Ctrl, RowMask>((uint32_t)(v >> 32)) << 32) | DppMove<Ctrl, RowMask>((uint32_t)v);
}

// Inclusive prefix sums over the 64 lanes (rows of 16 first, then the row
// totals carried by the two broadcasts).  Every lane must be active.
__device__ __forceinline__ uint32_t WaveInclusiveScan(uint32_t v) {
    v += DppMove<kDppShr1>(v);
    v += DppMove<kDppShr2>(v);
    v += DppMove<kDppShr4>(v);
    v += DppMove<kDppShr8>(v);
    v += DppMove<kDppBcast15, 0xa>(v);
    v += DppMove<kDppBcast31, 0xc>(v);
    return v;
}
__device__ __forceinline__ unsigned long long WaveInclusiveScan(unsigned long long v) {
    v += DppMove64<kDppShr1>(v);
    v += DppMove64<kDppShr2>(v);
    v += DppMove64<kDppShr4>(v);
    v += DppMove64<kDppShr8>(v);
    v += DppMove64<kDppBcast15, 0xa>(v);
    v += DppMove64<kDppBcast31, 0xc>(v);
    return v;
}

__device__ __forceinline__ unsigned long WaveInclusiveScan(unsigned long v) {
    return (unsigned long)WaveInclusiveScan((unsigned long long)v);
}

// Smallest value over the 64 lanes, in every lane's result (a min-scan,
// lane 63 read back).  Every lane must be active.
__device__ __forceinline__ uint32_t WaveMin(uint32_t v) {
    v = min(v, DppMove<kDppShr1>(v, ~0u));
    v = min(v, DppMove<kDppShr2>(v, ~0u));
    v = min(v, DppMove<kDppShr4>(v, ~0u));
    v = min(v, DppMove<kDppShr8>(v, ~0u));
    v = min(v, DppMove<kDppBcast15, 0xa>(v, ~0u));
    v = min(v, DppMove<kDppBcast31, 0xc>(v, ~0u));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Lane l's value as a scalar (v_readlane: read whatever the lane's
// register holds, active or not).
__device__ __forceinline__ uint32_t ReadLane(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ __forceinline__ uint64_t ReadLane64(uint64_t v, int l) {
    return ((uint64_t)ReadLane((uint32_t)(v >> 32), l) << 32) | ReadLane((uint32_t)v, l);
}

template <class T>
__device__ __forceinline__ T WaveBroadcast(T v, int src) {
    return __shfl(v, src, 64);
}

// Load through a pointer that came from memory (a table in a kernel
// argument, a job record, LDS), whose address space the compiler cannot
// infer: as a global load.  A flat load would also count against the LDS /
// scalar counter, so every wait for a shuffle or an LDS read would wait for
// the outstanding memory loads too.  Only for pointers into device memory.
template <class T>
__device__ __forceinline__ T LdG(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}

// Number of values below `a` among at most 64 sorted values held one per
// lane (lane i holds value i; lanes past the count hold ~0u), by six
// shuffles: no memory round trip.  At most 63 (a value above all 64 held
// ones reads as 63: callers compare the value there).  Every lane calls it.
__device__ __forceinline__ uint32_t RankInLanes(uint32_t held, uint32_t a) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1) {
        const uint32_t pv = __shfl(held, (int)(pos + step - 1), 64);
        if (pv < a) pos += step;
    }
    return pos;
}

__device__ __forceinline__ uint32_t Uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t Uniform64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// First index i in [0, n) with a[i] >= v (n if none).
__device__ __forceinline__ uint32_t LowerBound(const uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t m = (lo + hi) >> 1;
        if (a[m] < v) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__device__ __forceinline__ bool SortedContains(const uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
    uint32_t p = LowerBound(a, n, v);
    return p < n && a[p] == v;
}

}  // namespace cgd
