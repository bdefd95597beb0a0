// Host-side launchers of the gfx950 kernels (one translation unit each).
#pragma once
#include "engine.h"

namespace cgd {

// Small lists passed by value in the kernel argument segment.
struct EiList {
    uint32_t n;
    uint8_t ei[kMaxQE];
};
struct EiSeedInfo {
    uint32_t n;
    uint8_t ei[kMaxQE], i01[kMaxQE], i10[kMaxQE], vs0[kMaxQE], vs1[kMaxQE];
};

// index.hip
void LaunchDecode(const uint64_t* in, uint64_t* out, uint64_t n, uint32_t ushift, uint32_t umask, uint32_t ndir,
                  uint32_t vbits, hipStream_t s);
void LaunchUnique(const uint64_t* in, uint64_t* out, uint64_t n, uint32_t* flags, void* scan_scratch,
                  unsigned long long* n_out_dev, hipStream_t s);
void LaunchSegments(const uint64_t* keys, const unsigned long long* n_dev, uint64_t n_upper, uint32_t vbits,
                    uint32_t* flags, uint32_t* seg_row, uint64_t* seg_start, unsigned long long* n_segs_dev,
                    void* scan_scratch, hipStream_t s);
// Merges of sorted keys into (insert) or out of (remove) the lists they
// touch: one lane per short list, one wavefront per long one (queued in
// long_segs, n_upper entries; cnt->n_long must be zero at the launch).
void LaunchMergeInsert(const Tables& T, const uint64_t* keys, uint32_t vbits, const uint32_t* seg_row,
                       const uint64_t* seg_start, const unsigned long long* n_segs, uint64_t n_upper, uint32_t* sc_val,
                       uint32_t* sc_pos, uint32_t* long_segs, uint64_t pool_cap, Counters* cnt, hipStream_t s);
void LaunchMergeRemove(const Tables& T, const uint64_t* keys, uint32_t vbits, const uint32_t* seg_row,
                       const uint64_t* seg_start, const unsigned long long* n_segs, uint64_t n_upper,
                       uint32_t* long_segs, Counters* cnt, hipStream_t s);
void LaunchClearU(URec* u_rec, const uint32_t* u_rows, const unsigned long long* n_dev, uint64_t n_upper,
                  hipStream_t s);
void LaunchBuildU(URec* u_rec, const uint64_t* keys, uint64_t n, uint32_t vbits, uint32_t* vals,
                  const uint32_t* seg_row, const uint64_t* seg_start, const unsigned long long* n_segs, uint32_t* u_rows,
                  unsigned long long* u_rows_count, hipStream_t s);
void LaunchInitCaps(const uint64_t* seg_start, const unsigned long long* n_segs, uint64_t n_upper, uint64_t* caps,
                    hipStream_t s);
void LaunchInitLayout(const Tables& T, const uint64_t* keys, const unsigned long long* n_dev, uint64_t n_upper,
                      uint32_t vbits, const uint32_t* head_pos, const uint32_t* seg_row, const uint64_t* seg_start,
                      const unsigned long long* n_segs, const uint64_t* offs, Counters* cnt, hipStream_t s);
void LaunchRowLens(const Tables& T, uint64_t nrows, int which, uint64_t* lens, hipStream_t s);
void LaunchExport(const Tables& T, uint64_t nrows, int which, const uint64_t* row_off, uint32_t ushift, uint64_t* out,
                  hipStream_t s);

// masks.hip
void LaunchMaskClear(uint64_t* masks, uint64_t ei_words, const EiList& L, hipStream_t s);
void LaunchMaskSeed(const Tables& T, const uint32_t* u_rows, const unsigned long long* n_dev, uint64_t n_upper,
                    const EiSeedInfo& info, hipStream_t s);
void LaunchMaskDepth(const Tables& T, uint32_t d, const EiList& L, Counters* cnt, hipStream_t s);

// match.hip
void LaunchSeed(const Tables& T, const EiList& L, uint32_t* rows, uint64_t cap_rows, Counters* cnt, hipStream_t s);
// Rows of `words` u32 (1 or 2; word 0 = tag << tag_shift | v0) of `in`
// copied to `out` grouped by tag; hist = 2 * kMaxQE zeroed counters.
void LaunchGroupByTag(const uint32_t* in, uint64_t n, uint32_t words, uint32_t tag_shift, unsigned long long* hist,
                      uint32_t* out, hipStream_t s);
// Depth-1 rows [(ei << tag_shift) | v0] (one word each).
void LaunchSeed1(const Tables& T, const EiList& L, uint32_t* rows, uint64_t cap_rows, Counters* cnt, hipStream_t s);
void LaunchExtend(const Tables& T, const uint32_t* in, uint64_t n_in, uint32_t depth, uint32_t* out,
                  uint64_t cap_rows, Counters* cnt, hipStream_t s);
// One query's share of a final-DFS launch: its index tables, its frontier
// rows (depth D, word 0 = ei tag | v0), its counters (per-edge counts,
// duplicate flag of its update lists) and its counting options.
enum : uint32_t {
    kJobShortest = 1u,  // scan the shortest backward-neighbour list when lists are sets
    kJobSuffix = 2u,    // count the product suffix of the order in closed form
    kJobWaveSearch = 4u,  // search long backward-neighbour lists with the whole wavefront
    kJobWaveShift = 8u,   // bits 8..12: log2 of the shortest list searched so (0: 256)
};
struct DfsJob {
    Tables T;
    const uint32_t* rows;
    Counters* cnt;
    uint32_t D;
    uint32_t opts;  // kJob* bits
};
// Final DFS over the rows of njobs queries (row_base[j] = first global row of
// job j, row_base[njobs] = all rows); `cnt` holds the shared work queue.
// surv_levels >= qv - 1 of every job (survivor stacks in LDS); blocks_per_cu
// 0 = as many as fit.
// cache_cap: u32 per wave of the stable-intersection cache (0: all the LDS
// left; smaller values leave part of it unused).
void LaunchDfs(const DfsJob* jobs, uint32_t njobs, const unsigned long long* row_base, uint64_t n_rows, Counters* cnt,
               const StealQueue& Q, uint32_t* tuples, uint64_t tup_cap, uint32_t num_cus, uint32_t blocks_per_cu,
               uint32_t grab, uint32_t surv_levels, uint32_t cache_cap, hipStream_t s);
// LDS plan of a final-DFS launch (see DfsLdsShapeT in match.hip).
struct DfsShape {
    uint32_t blocks_per_cu, dyn_bytes, stab_cap;
};
DfsShape DfsLdsShape(uint32_t surv_levels, bool tuples);
// Resident KDfs blocks per CU at full occupancy.
uint32_t DfsBlocksPerCU();

void LaunchCountDups(const uint64_t* k, uint64_t n, Counters* cnt, hipStream_t s);

void LaunchQueueSync(Counters* cnt, hipStream_t s);

// ---------------------------------------------------------------- GPMA layout
constexpr uint32_t kPmaLeaf = 512;   // slots per leaf (u64 keys)
constexpr uint32_t kPmaVBits = 27;   // PMA key = (row << 27) | v
struct PmaWin {                       // a window being rebalanced
    uint32_t a, nl;                   // first leaf, leaves
    uint64_t E, ilo, ihi;             // existing keys, insert range [ilo, ihi)
};
void LaunchPmaEncode(const uint64_t* in, uint64_t n, uint32_t vbits, uint64_t* out, hipStream_t s);
void LaunchPmaSpread(const uint64_t* keys, uint64_t n, uint64_t* slot, uint32_t* cnt, uint64_t* lo, uint32_t nleaves,
                     hipStream_t s);
void LaunchPmaRoute(const uint64_t* keys, uint64_t n, const uint64_t* lo, uint32_t nleaves, uint32_t* leaf_of,
                    hipStream_t s);
void LaunchPmaAbsent(const uint64_t* keys, uint64_t n, const uint32_t* leaf_of, const uint64_t* slot,
                     const uint32_t* cnt, uint32_t* keep, hipStream_t s);
void LaunchPmaCompact(const uint64_t* keys, const uint32_t* leaf_of, uint64_t n, const uint32_t* keep,
                      const uint32_t* pos, uint64_t* out_keys, uint32_t* out_leaf, hipStream_t s);
void LaunchPmaWindows(const uint32_t* cnt, uint32_t* ins_cnt, const uint32_t* ins_leaf,
                      const unsigned long long* n_ins_dev, uint64_t n_upper, uint64_t* S, uint8_t* chosen,
                      uint32_t nleaves, uint32_t height, Counters* counters, hipStream_t s);
void LaunchPmaRebalance(const uint8_t* chosen, uint32_t nleaves, uint32_t height, uint32_t* wins,
                        unsigned long long* nwin, uint32_t* win_of, const uint32_t* gpos, const uint32_t* ins_leaf,
                        const uint64_t* ins_keys, const unsigned long long* n_ins_dev, uint64_t n_upper, PmaWin* meta,
                        uint64_t* slot, uint64_t* shadow, uint32_t* cnt, uint64_t* lo, unsigned long long* level_seen,
                        hipStream_t s);
void LaunchPmaDelete(const uint64_t* dkeys, const uint32_t* dleaf, uint64_t n, uint64_t* slot, uint32_t* cnt,
                     hipStream_t s);
void LaunchPmaFlat(const uint64_t* slot, const uint32_t* cnt, const uint32_t* gpos, uint32_t nleaves, uint64_t* flat,
                   hipStream_t s);
void LaunchPmaRows(const uint64_t* flat, uint64_t n, uint32_t* pool, MRec* m_rec, uint64_t rows, hipStream_t s);

}  // namespace cgd
