// 1-bit support masks on gfx950 (BuildLocalIndexBitAll,
// graph/graph_gpu.cu:803-869; kernels setInitialValidBitsAll1Bit and
// checkAllConstraintsAndSetBitsAll1Bit, kernels/indexing.cu:2295-2379).
//
// Tile (ei, qv) holds one bit per compressed vertex in u64 words (a wave64
// ballot is exactly one word).  Tiles of vertices that are not rebuilt for ei
// are never stored: DevPlan::all_ones makes every read return 1.
//   depth 0/1: sparse -- only owners of this batch's U lists get a bit, so
//              the kernel walks U's touched rows instead of all vertices.
//   depth >= 2: dense -- one wavefront per mask word; each lane checks its
//              vertex against every backward neighbour's list (M, plus U when
//              that query edge's index is <= ei).
#include "engine.h"
#include "kernels.h"
#include "wave.h"

namespace cgd {

namespace {

__device__ __forceinline__ bool TileBitM(const Tables& T, uint32_t ei, uint32_t qv, uint32_t v) {
    if (T.P->all_ones[ei][qv]) return true;
    return (T.masks[((uint64_t)ei * kMaxQV + qv) * T.W + (v >> 6)] >> (v & 63)) & 1ull;
}

__global__ void KMaskClear(uint64_t* __restrict__ masks, uint64_t ei_words, EiList L) {
    const uint32_t ei = L.ei[blockIdx.y];
    uint64_t* base = masks + (uint64_t)ei * ei_words;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ei_words; i += (uint64_t)gridDim.x * blockDim.x)
        base[i] = 0;
}

__global__ void KMaskSeed(Tables T, const uint32_t* __restrict__ u_rows, const unsigned long long* __restrict__ n_dev,
                          EiSeedInfo info) {
    const uint64_t n = *n_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t row = u_rows[i];
        const uint32_t e = row % T.ndir, u = row / T.ndir;
        const uint64_t bit = 1ull << (u & 63);
        for (uint32_t k = 0; k < info.n; k++) {
            const uint32_t ei = info.ei[k];
            if (e == info.i01[k])
                atomicOr((unsigned long long*)&T.masks[((uint64_t)ei * kMaxQV + info.vs0[k]) * T.W + (u >> 6)], bit);
            if (e == info.i10[k])
                atomicOr((unsigned long long*)&T.masks[((uint64_t)ei * kMaxQV + info.vs1[k]) * T.W + (u >> 6)], bit);
        }
    }
}

// Algorithmic bytes counted into cnt->bytes: 4 per list entry read and 4 per
// (backward neighbour, vertex) list size read; the tile write is added by
// the host (ceil(dv/8) per rebuilt tile).
__global__ void __launch_bounds__(256) KMaskDepth(Tables T, uint32_t d, EiList L, Counters* __restrict__ cnt) {
    const uint32_t ei = L.ei[blockIdx.y];
    const cgbdsm_plan& p = T.P->p;
    const uint32_t u0 = p.io_vs[ei][d];
    const uint32_t bs = p.io_bni_offs[ei][d], be = p.io_bni_offs[ei][d + 1];
    const uint32_t lane = LaneId();
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / kWave);
    uint64_t* out = T.masks + ((uint64_t)ei * kMaxQV + u0) * T.W;
    unsigned long long reads = 0;
    for (uint64_t w = (uint64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; w < T.W; w += nw) {
        const uint32_t v = (uint32_t)(w * 64 + lane);
        bool ok = v < T.dv;
        for (uint32_t k = bs; k < be; k++) {
            const uint32_t u1 = p.io_vs[ei][p.io_bni[ei][k]];
            const uint32_t bwd = p.eidx[u0 * kMaxQV + u1];
            const bool up_vis = p.dir_to_edge[bwd] <= ei;
            bool sat = false;
            if (ok) {
                const uint64_t r = (uint64_t)v * T.ndir + bwd;
                const MRec m = T.m_rec[r];
                const uint32_t* L0 = T.pool + m.off;
                uint32_t j = 0;
                for (; j < m.len && !sat; j++) sat = TileBitM(T, ei, u1, L0[j]);
                reads += 1 + j;
                if (!sat && up_vis) {
                    const URec uu = T.u_rec[r];
                    const uint32_t* L1 = T.u_vals + uu.off;
                    for (j = 0; j < uu.len && !sat; j++) sat = TileBitM(T, ei, u1, L1[j]);
                    reads += 1 + j;
                }
            }
            ok = ok && sat;
            if (!__ballot(ok)) break;
        }
        const uint64_t word = __ballot(ok);
        if (lane == 0) out[w] = word;
    }
    const unsigned long long tot = ReadLane64(WaveInclusiveScan(reads), 63);
    if (lane == 0 && tot) atomicAdd(&cnt->bytes, 4ull * tot);
}

}  // namespace

void LaunchMaskClear(uint64_t* masks, uint64_t ei_words, const EiList& L, hipStream_t s) {
    if (!L.n || !ei_words) return;
    unsigned gx = (unsigned)std::min<uint64_t>((ei_words + 255) / 256, 1024);
    KMaskClear<<<dim3(gx, L.n), 256, 0, s>>>(masks, ei_words, L);
}

void LaunchMaskSeed(const Tables& T, const uint32_t* u_rows, const unsigned long long* n_dev, uint64_t n_upper,
                    const EiSeedInfo& info, hipStream_t s) {
    if (!n_upper || !info.n) return;
    unsigned g = (unsigned)std::min<uint64_t>((n_upper + 255) / 256, 4096);
    KMaskSeed<<<g, 256, 0, s>>>(T, u_rows, n_dev, info);
}

void LaunchMaskDepth(const Tables& T, uint32_t d, const EiList& L, Counters* cnt, hipStream_t s) {
    if (!L.n || !T.W) return;
    unsigned gx = (unsigned)std::min<uint64_t>((T.W + 3) / 4, 4096);
    KMaskDepth<<<dim3(gx, L.n), 256, 0, s>>>(T, d, L, cnt);
}

}  // namespace cgd
