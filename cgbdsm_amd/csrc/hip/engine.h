// Internal interface of the gfx950 Stage-2 engine.
//
// Device layout of one query's index (all in HBM, sized for 288 GB parts):
//   row r = owner * ndir + dir            (owner-major: DV growth only appends)
//   MRec m_rec[dv_cap * ndir]             main index M: {u64 off, u32 len, u32 cap}
//   u32  pool[pool_cap]                   sorted target lists of M, with slack
//   URec u_rec[dv_cap * ndir]             update index U of the current batch:
//                                         {u32 off, u32 len} into u_vals
//   u64  masks[qe][12][W]                 1-bit support masks, W = ceil(dv/64)
//   u32  frontier[rows][depth]            partial matches, word 0 = ei tag | v0
// Keys are re-encoded on the device as (row << vbits) | v, vbits = bits(dv-1),
// so one LSD radix sort groups them by list and orders each list.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "cgbdsm.h"

namespace cgd {

constexpr int kMaxQV = CGBDSM_MAX_QV;
constexpr int kMaxQE = CGBDSM_MAX_QE;
constexpr int kMaxDir = CGBDSM_MAX_DIR;
constexpr int kWave = 64;

struct alignas(16) MRec {
    uint64_t off;
    uint32_t len;
    uint32_t cap;
};
struct alignas(8) URec {
    uint32_t off;
    uint32_t len;
};

// Device counters block (one per query), zeroed by a fill kernel.
struct alignas(128) Counters {
    unsigned long long n_unique;      // keys after unique
    unsigned long long n_segs;        // segments (touched lists)
    unsigned long long pool_top;      // next free pool slot
    unsigned long long rows_out;      // frontier rows written by the last launch
    unsigned long long u_dups;        // adjacent equal keys in the sorted update keys
    unsigned long long flags;         // bit 0: pool overflow, bit 1: frontier overflow, bit 2: tuple overflow
    unsigned long long bytes;         // algorithmic bytes of the last index pass
    unsigned long long rows_touched;  // lists rewritten by the last index pass
    unsigned long long u_rows;        // touched rows of the current U
    unsigned long long step_lanes;    // candidates covered by those steps (active lanes)
    unsigned long long batched;       // steps whose product suffixes were counted together
    unsigned long long count_ovf;     // final DFS: a count did not fit 64 bits (precedes per_edge)
    unsigned long long per_edge[kMaxQE];
    unsigned long long card[kMaxDir], nodes[kMaxDir];
    unsigned long long tuples;
    unsigned long long steps;         // 64-lane candidate scan steps of the final DFS
    unsigned long long scanned;       // candidate list entries read by the final DFS
    unsigned long long probes;        // binary-search probes of the final DFS
    unsigned long long donated;       // partial matches handed to the stealing queue
    unsigned long long pma_level;     // GPMA: highest rebalanced window level + 1 (since reset)
    unsigned long long wave_ticks;    // final DFS: summed wave lifetimes (wall-clock ticks)
    unsigned long long idle_ticks;    // final DFS: summed time waves waited for work
    unsigned long long n_long;        // index merges: touched lists left to one wavefront each
    // Words polled by many wavefronts at once, one 128-byte L2 line each.
    struct alignas(128) Line {
        unsigned long long v;
        unsigned long long pad[15];
    };
    Line q_head, q_tail, q_pending, q_idle, q_rows;
};

enum : unsigned long long {
    kFlagPool = 1,
    kFlagFrontier = 2,
    kFlagTuples = 4,
    kFlagSteal = 8,
    kFlagAborted = 16,
    kFlagBadRow = 32   // a final-DFS frontier row held an out-of-range tag or vertex id
};

// One stolen piece of work: positions 0..depth-1 of the global order are
// fixed, and candidates [lo, hi) of position `depth` remain to be scanned.
struct alignas(16) StealItem {
    uint32_t depth, ei;
    uint32_t v[CGBDSM_MAX_QV];
    uint32_t lo, hi;
};

struct StealQueue {
    StealItem* items;
    unsigned long long* seq;  // seq[slot] = ticket + 1 once the item is published
    uint64_t cap;             // slots (power of two)
    uint32_t donate_min;      // remaining survivors needed before donating (0: never)
    uint32_t range_min;       // candidates a donated range must at least hold
    const uint32_t* abort;    // host-mapped coherent flag, non-zero asks kernels to stop
    long long poll_ticks;     // wall-clock ticks between a wave's reads of shared queue words
    long long host_ticks;     // wall-clock ticks between a wave's reads of the host flag
};

// Per-query parameter block in device memory (read through scalar loads).
// Replaces the __constant__ tables of graph/graph_gpu.cu:35-46 and
// utils/globals.cu:8-16, one block per registered query.
struct DevPlan {
    cgbdsm_plan p;
    uint8_t all_ones[kMaxQE][kMaxQV];  // tile is not rebuilt: every bit reads 1
    uint8_t sa[kMaxQE], sb[kMaxQE];    // the two directions of edge ei
    uint32_t tag_shift;                // frontier word 0 = ei << tag_shift | v0
    uint32_t v_mask;                   // (1 << tag_shift) - 1
    uint32_t cp_start;                 // first position of the product suffix (qv: none)
    uint32_t stab_pos;                 // suffix position whose stable lists are cached (0xFF: none)
    uint32_t stab_var_k;               // its backward neighbour at the deepest position
    uint32_t stab_deps;                // positions the cached intersection depends on
    uint32_t sfx_batch;                // every label class of the product suffix has one member:
                                       // the DFS counts the suffixes of a whole step of
                                       // siblings at the position before it at once
    uint8_t maxbn[kMaxQV];             // largest backward-neighbour position of each position
    uint32_t dep_mask[kMaxQV];         // positions whose candidate set depends on position x
};

// Device pointers of one query's index, passed by value to kernels.
struct Tables {
    const DevPlan* P;
    MRec* m_rec;
    URec* u_rec;
    uint32_t* pool;
    uint32_t* u_vals;
    uint64_t* masks;
    uint32_t ndir, qv, qe, dv, W;
};

__host__ __device__ inline uint32_t CapFor(uint32_t len) {
    uint32_t want = len + (len >> 1);
    if (want >= (1u << 30)) return (want + 63u) & ~63u;
    uint32_t c = 8;
    while (c < want) c <<= 1;
    return c;
}

struct Status {
    cgbdsm_status code = CGBDSM_OK;
    std::string msg;
};

#define CGD_HIP(x)                                                                     \
    do {                                                                               \
        hipError_t _e = (x);                                                           \
        if (_e != hipSuccess) throw ::cgd::HipError(_e, #x, __FILE__, __LINE__);       \
    } while (0)

struct HipError {
    hipError_t e;
    std::string what;
    HipError(hipError_t err, const char* expr, const char* file, int line) : e(err) {
        char buf[512];
        snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", expr, hipGetErrorString(err), file, line);
        what = buf;
    }
};
struct ApiError {
    cgbdsm_status code;
    std::string what;
};

inline uint32_t BitsFor(uint64_t x) {  // bits needed to hold values < x (x >= 1)
    uint32_t b = 1;
    while (b < 64 && (1ull << b) < x) b++;
    return b;
}

// ---------------------------------------------------------------- primitives
// Exclusive scan of n u32 (or u64) values into out (may alias in); *total gets the sum.
void ScanU32(const uint32_t* in, uint32_t* out, uint64_t n, uint64_t* total_dev, void* scratch, hipStream_t s);
void ScanU64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* total_dev, void* scratch, hipStream_t s);
size_t ScanScratchBytes(uint64_t n);
// LSD radix sort of n u64 keys over the low `bits` bits; result in keys (alt is scratch).
void RadixSortU64(uint64_t* keys, uint64_t* alt, uint64_t n, uint32_t bits, uint32_t* hist, hipStream_t s);
size_t RadixHistWords(uint64_t n);
void FillU64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);
// Up to 16 words of one block set in one launch (word offsets from base).
struct WordSet {
    uint32_t n;
    uint32_t off[16];
    uint64_t val[16];
};
void SetWords(uint64_t* base, const WordSet& ws, hipStream_t s);
void FillU32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);

}  // namespace cgd
