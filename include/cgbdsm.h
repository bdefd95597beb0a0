/*
 * cgbdsm.h -- C ABI of the MI355X (gfx950) Stage-2 engine for continuous
 * subgraph matching (CGBDSM's GPU stage).
 *
 * The host pipeline (Stage-1 CaLiG producer, planner, cgcsm_* CLI) calls the
 * device only through these functions.  Arguments are plain pointers and sizes;
 * every function returns a cgbdsm_status and never exits the process.
 *
 * Each entry point names the reference interface it replaces (file:line in the
 * CGBDSM sources):
 *   CaLiGHelperGamma  index/calig_helper_gamma.h:91-110
 *   GPUGraphLoader    graph/graph_gpu.h:152-188
 *
 * Key format at the boundary (same as index/calig_helper_gamma.cu:264):
 *   key = (dir_edge << (64 - ebits)) | (u << 32) | v
 * with ebits = 5 (reference-compatible: at most 32 directed query edges,
 * 27-bit owner ids) or ebits = 7 (wide: 84 directed edges, 25-bit owner ids).
 */
#ifndef CGBDSM_H
#define CGBDSM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compile-time limits of the query model (utils/config.h:5-6). */
#define CGBDSM_MAX_QV 12
#define CGBDSM_MAX_QE 42
#define CGBDSM_MAX_DIR (2 * CGBDSM_MAX_QE)
#define CGBDSM_MAX_BATCH_QUERIES 256
#define CGBDSM_NO_EDGE 0xFFu

typedef enum {
    CGBDSM_OK = 0,
    CGBDSM_ERR_INVALID = 1,   /* bad argument or call order               */
    CGBDSM_ERR_OOM = 2,       /* device memory budget exceeded            */
    CGBDSM_ERR_DEVICE = 3,    /* HIP runtime error                        */
    CGBDSM_ERR_LIMIT = 4,     /* query or key exceeds an encoding limit   */
    CGBDSM_ERR_OVERFLOW = 5   /* a bounded device queue or spin overflowed */
} cgbdsm_status;

/* Cartesian-product tags of the global order (graph/plan.h:43-49). */
enum {
    CGBDSM_CP_NONE = 0,
    CGBDSM_CP_TREE_SINGLE = 1,
    CGBDSM_CP_NONTREE_SINGLE = 2,
    CGBDSM_CP_TREE_PRODUCT = 3,
    CGBDSM_CP_NONTREE_PRODUCT = 4
};

/*
 * Query model and plan tables: the POD form of QueryGraph (graph/graph.h:18-46),
 * OrderPerEdge (graph/plan.h:12-18) and the Plan flags (graph/plan.h:50-59).
 * These replace the __constant__ symbols of graph/graph_gpu.cu:35-46 and
 * utils/globals.cu:8-16 with one per-query parameter block, so several
 * queries can live in one process.
 */
typedef struct {
    uint32_t qv;                        /* query vertices (3..12)              */
    uint32_t qe;                        /* undirected query edges (2..42)      */
    uint32_t ndir;                      /* directed query edges = 2*qe         */
    uint32_t ebits;                     /* 5 (reference keys) or 7 (wide keys) */
    uint8_t eidx[CGBDSM_MAX_QV * CGBDSM_MAX_QV];   /* [u*12+uu] -> dir edge or 0xFF */
    uint8_t dir_to_edge[CGBDSM_MAX_DIR];           /* dir edge -> undirected edge   */
    uint8_t qe_u[CGBDSM_MAX_QE];                   /* qe_list_[i].first  (u < uu)   */
    uint8_t qe_v[CGBDSM_MAX_QE];                   /* qe_list_[i].second            */
    /* global matching order (Plan::global_order_) */
    uint8_t go_vs[CGBDSM_MAX_QV];
    uint8_t go_bni_offs[CGBDSM_MAX_QV + 1];
    uint8_t go_bni[CGBDSM_MAX_DIR];                /* positions, sorted by selectivity */
    uint8_t go_cp[CGBDSM_MAX_QV];                  /* global_cartesian_product_info_   */
    uint8_t go_lcls[CGBDSM_MAX_QV];                /* label class of each position: the
                                                      first position with the same query
                                                      vertex label (no reference field) */
    /* per-edge indexing orders (Plan::indexing_orders_) */
    uint8_t io_vs[CGBDSM_MAX_QE][CGBDSM_MAX_QV];
    uint8_t io_bni_offs[CGBDSM_MAX_QE][CGBDSM_MAX_QV + 1];
    uint8_t io_bni[CGBDSM_MAX_QE][CGBDSM_MAX_DIR];
    uint8_t rebuild_v[CGBDSM_MAX_QE][CGBDSM_MAX_QV];  /* Plan::rebuild_v_flags_ */
} cgbdsm_plan;

/* Per-batch device statistics (filled by cgbdsm_stats). */
typedef struct {
    uint64_t batches;
    uint64_t dv_count;            /* compressed data vertices                   */
    uint64_t main_entries;        /* live entries in the main index M           */
    uint64_t pool_used;           /* u32 slots taken from the list pool         */
    uint64_t pool_cap;
    uint64_t last_ins, last_del, last_upd;   /* keys staged for the last batch  */
    uint64_t last_rows_touched;   /* (edge,vertex) lists rewritten, last batch  */
    uint64_t last_idx_bytes;      /* algorithmic index-maintenance bytes        */
    uint64_t last_frontier_rows;  /* rows handed to the final DFS               */
    uint64_t last_frontier_depth;
    uint64_t last_matches;
    uint64_t compactions;
    uint64_t last_dfs_scanned;    /* candidate list entries read by the final DFS */
    uint64_t last_dfs_probes;     /* binary-search probes of the final DFS        */
    uint64_t last_dfs_donated;    /* work items handed between DFS wavefronts     */
    uint64_t gpma_rebuilds;       /* GPMA layout: growths by full rebuild         */
    uint64_t gpma_windows;        /* GPMA layout: windows rebalanced              */
    uint64_t gpma_top_level;      /* GPMA layout: highest window level + 1        */
    uint64_t last_dfs_steps;      /* 64-lane candidate steps of the final DFS's
                                     descent loop (suffix sizing excluded)        */
    uint64_t last_dfs_step_lanes; /* candidates those steps covered: lane
                                     utilisation = step_lanes / (64 * steps)      */
    uint64_t last_dfs_batched;    /* steps whose product suffixes were counted
                                     for all their survivors at once              */
    uint64_t last_dfs_lds_cache;  /* u32 per wave of the stable-intersection
                                     cache in the last final-DFS launch           */
    double   ms_apply, ms_update_index, ms_masks, ms_match, ms_commit; /* last batch, device time */
    double   last_dfs_busy;       /* share of final-DFS wave time spent working   */
} cgbdsm_stats;

typedef struct cgbdsm_ctx cgbdsm_ctx;
typedef struct cgbdsm_query cgbdsm_query;

/* Most recent error text of the calling thread ("" if none). */
const char* cgbdsm_last_error(void);

/* Library build identification (gfx target, key widths). */
const char* cgbdsm_version(void);

/* Device context: one HIP device, one stream, a memory budget.
 * Replaces cudaSetDevice + the MemPool of GPUGraphLoader's ctor
 * (main_calig.cpp:64, graph/graph_gpu.cu:58-102). budget_bytes = 0: no cap. */
cgbdsm_status cgbdsm_ctx_create(int device, uint64_t budget_bytes, cgbdsm_ctx** out);
cgbdsm_status cgbdsm_ctx_destroy(cgbdsm_ctx* ctx);
/* Device bytes held by the context's queries and its budget (0: none); a
 * serving loop admits work against them.  No reference counterpart (its
 * pool is a fixed 256 MB, utils/config.h:19). */
cgbdsm_status cgbdsm_ctx_memory(cgbdsm_ctx* ctx, uint64_t* used_bytes, uint64_t* budget_bytes);

/* Register a query: GPUGraphLoader ctor + LoadQuery + AllocRelations
 * (graph/graph_gpu.h:152-169, graph/graph_gpu.cu:156-197).  Only the query
 * model fields of `plan` (qv, qe, ndir, ebits, eidx, dir_to_edge, qe_u, qe_v,
 * io_*) are read here; the matching order arrives with cgbdsm_query_set_plan. */
cgbdsm_status cgbdsm_query_create(cgbdsm_ctx* ctx, const cgbdsm_plan* plan, cgbdsm_query** out);
cgbdsm_status cgbdsm_query_destroy(cgbdsm_query* q);

/* Upload the global matching order, rebuild flags and cartesian tags:
 * GPUGraphLoader::LoadPlan (graph/graph_gpu.h:159, graph/graph_gpu.cu:174-193). */
cgbdsm_status cgbdsm_query_set_plan(cgbdsm_query* q, const cgbdsm_plan* plan);

/* Storage layout of the main index, chosen before cgbdsm_index_init.
 * GAMMA: per-(edge, vertex) sorted lists with slack, merged in place
 *   (index/calig_helper_gamma.cu, kernels/indexing.cu:296-458).
 * GPMA: a packed-memory array of all index keys with density-bounded
 *   rebalancing (index/gpma_wrapper.cu:19-116, gpma/gpma.cu), extracted into
 *   the same flat lists for matching every batch
 *   (index/calig_helper_gpma.cu:383-414).  Both yield identical lists. */
enum { CGBDSM_LAYOUT_GAMMA = 0, CGBDSM_LAYOUT_GPMA = 1 };
cgbdsm_status cgbdsm_query_set_layout(cgbdsm_query* q, int layout);

/* Initial main index from host keys (any order, duplicates allowed):
 * CaLiGHelperGamma::ConvertGlobalIndex device half
 * (index/calig_helper_gamma.h:95, index/calig_helper_gamma.cu:238-360).
 * Writes cardinalities[ndir] (entries per directed edge) and
 * avg_degrees[ndir] (entries / non-empty lists) when non-NULL. */
cgbdsm_status cgbdsm_index_init(cgbdsm_query* q, const uint64_t* keys, uint64_t n,
                                uint32_t dv_count, uint32_t* cardinalities, float* avg_degrees);

/* Hand one batch's host key vectors to the device (copied before return).
 * Replaces the staging vectors el_host_stage_ / del_host_stage_ /
 * el_update_host_stage_ and edge_ok_per_batch_ (index/calig_helper_gamma.h:60-74)
 * plus AllocateFromMappingGPU (index/calig_helper_gamma.cu:107-130). */
cgbdsm_status cgbdsm_batch_stage(cgbdsm_query* q,
                                 const uint64_t* ins, uint64_t n_ins,
                                 const uint64_t* del, uint64_t n_del,
                                 const uint64_t* upd, uint64_t n_upd,
                                 const uint8_t* edge_ok, uint32_t dv_count);

/* Device-resident batch FIFO.  cgbdsm_batch_enqueue validates and copies a
 * batch's host keys into HBM (the boundary copy); cgbdsm_batch_stage_next
 * makes the oldest enqueued batch the staged one without any copy.  Lets the
 * Stage-1 thread run ahead of Stage 2 (main_calig.cpp:203-289) and lets
 * callers time Stage 2 with its inputs already resident. */
cgbdsm_status cgbdsm_batch_enqueue(cgbdsm_query* q,
                                   const uint64_t* ins, uint64_t n_ins,
                                   const uint64_t* del, uint64_t n_del,
                                   const uint64_t* upd, uint64_t n_upd,
                                   const uint8_t* edge_ok, uint32_t dv_count);
cgbdsm_status cgbdsm_batch_stage_next(cgbdsm_query* q, int* staged);

/* M <- (M u Ins) \ Del: CaLiGHelperGamma::TransferUpdateAll
 * (index/calig_helper_gamma.h:105, index/calig_helper_gamma.cu:485-517). */
cgbdsm_status cgbdsm_transfer_update_all(cgbdsm_query* q);

/* U <- sorted Upd lists: CaLiGHelperGamma::TransferUpdateIndex
 * (index/calig_helper_gamma.h:106, index/calig_helper_gamma.cu:569-597).
 * *nonempty = 0 when the batch has no update keys (the caller skips it,
 * main_calig.cpp:301-302). */
cgbdsm_status cgbdsm_transfer_update_index(cgbdsm_query* q, int* nonempty);

/* Per-(query edge, query vertex) 1-bit support masks:
 * GPUGraphLoader::BuildLocalIndexBitAll (graph/graph_gpu.h:173, graph/graph_gpu.cu:803-869). */
cgbdsm_status cgbdsm_build_local_index(cgbdsm_query* q);

/* Merged matching of all query edges: GPUGraphLoader::MatchingBitAll
 * (graph/graph_gpu.h:180, graph/graph_gpu.cu:1261-1480).  Adds the batch's
 * positive matches to *num_matches; per_edge[qe] (optional) receives the
 * count attributed to each query edge (the 5-bit edge tag).  May be called
 * again before cgbdsm_batch_update_global_index: it counts the same batch
 * again (the index is only read), e.g. to time other launch shapes. */
cgbdsm_status cgbdsm_matching(cgbdsm_query* q, unsigned long long* num_matches, uint64_t* per_edge);

/* Optional first half of cgbdsm_matching: the depth-2 seeds and the BFS
 * levels (writeInitialResultsAllBit and the write_res levels of
 * extendBFSAllBit, graph/graph_gpu.cu:1261-1340) on the query's stream,
 * leaving the frontier for the final DFS of the next cgbdsm_matching or
 * cgbdsm_matching_batch call.  Lets a serving loop build the frontiers of
 * several queries concurrently (one host thread per query) before one
 * batched DFS launch.  Its last kernels may still run on the query's
 * stream on return; cgbdsm_matching_batch orders its launch after them. */
cgbdsm_status cgbdsm_build_frontier(cgbdsm_query* q);

/* cgbdsm_matching for nq registered queries of one context at once: each
 * query's seeds and BFS levels run on its own stream (unless
 * cgbdsm_build_frontier built them already), then ONE persistent
 * final-DFS launch counts all of them from a shared device work queue (the
 * first query's), so that the queries' uneven tails fill each other's idle
 * compute units.  Adds query i's matches to num_matches[i]; per_edge[i]
 * (optional array, entries optional) receives its per-edge counts.  Counts
 * are identical to nq separate cgbdsm_matching calls.  Every query must be
 * in the state cgbdsm_matching expects; the deadline and the statistics of
 * the launch are the first query's (last_dfs_* are stored in each query).
 * No reference counterpart: the reference runs one query per process
 * (main_calig.cpp:29-35).  nq <= CGBDSM_MAX_BATCH_QUERIES. */
cgbdsm_status cgbdsm_matching_batch(cgbdsm_query* const* qs, int nq, unsigned long long* num_matches,
                                    uint64_t* const* per_edge);

/* Post-match M <- M u Upd: CaLiGHelperGamma::BatchUpdateGlobalIndex
 * (index/calig_helper_gamma.h:98, index/calig_helper_gamma.cu:539-567). */
cgbdsm_status cgbdsm_batch_update_global_index(cgbdsm_query* q);

/* Matching knobs (no reference counterpart beyond utils/config.h:22):
 * bfs_row_limit = MIN_NRESULTS_TO_GPU analogue (rows below which a BFS
 * level is expanded before the final DFS); frontier_cap = capacity of the
 * frontier buffer in rows (small values force the overflow fallback of
 * graph/graph_gpu.cu:1323-1327); donate_min = pending candidates at which a
 * DFS wavefront donates work to the global stealing queue (0 disables). */
cgbdsm_status cgbdsm_set_match_params(cgbdsm_query* q, uint64_t bfs_row_limit,
                                      uint64_t frontier_cap, uint32_t donate_min);

/* Final-DFS refinements, both on by default; counts are identical either
 * way (no reference counterpart: the reference scans the first backward
 * neighbour's list, kernels/enumeration.cu:884-1000, and enumerates its
 * cartesian products, kernels/cartesian_product.cu:264-427).
 * shortest_list: each position scans its shortest backward-neighbour list
 *   when the batch's update lists hold no duplicate value.
 * product_suffix: the tail of the order whose positions are mutually
 *   non-adjacent is counted in closed form (products of candidate-set
 *   sizes, inclusion-exclusion within a label class) once per partial
 *   match of the positions before it. */
cgbdsm_status cgbdsm_set_dfs_options(cgbdsm_query* q, int shortest_list, int product_suffix);

/* Launch shape of the final DFS (autotuner knobs; counts are identical for
 * every setting).  rows_per_grab = frontier rows a wavefront takes per
 * cursor atomic (0: default 4, at most 1024); blocks_per_cu = resident
 * blocks per compute unit (0: as many as fit; larger values are clamped).
 * For cgbdsm_matching_batch the first query's setting applies.  No
 * reference counterpart (fixed GRID_DIM/BLOCK_DIM, utils/config.h). */
cgbdsm_status cgbdsm_set_dfs_tuning(cgbdsm_query* q, uint32_t rows_per_grab, uint32_t blocks_per_cu);

/* Final-DFS list search (counts are identical either way; no reference
 * counterpart).  On, every membership test of the lanes' candidates in a
 * list the whole wavefront shares is done by all lanes at once: a list of
 * at most 64 entries is loaded whole (one entry per lane) and searched by
 * shuffles; a list of wave_search entries or more is cut into 64 buckets
 * (each lane loads one pivot, each candidate binary-searches one bucket);
 * lists in between are binary-searched per lane.  wave_search = a power of
 * two >= 64 (the default is 64), 1 = 256, 0 = off (one binary search per
 * lane and list). */
cgbdsm_status cgbdsm_set_dfs_search(cgbdsm_query* q, int wave_search);

/* Stealing chunk and LDS split of the final DFS (autotuner knobs; counts
 * are identical for every setting).  range_min = smallest candidate range a
 * wavefront splits off for idle wavefronts (0: default 1024; at least 64;
 * the reference's fixed STEAL_CHUNK / TASK_SHARE_BOUND,
 * kernels/enumeration_balance.cu:11-17); cache_cap = u32 per wavefront of
 * the stable-intersection cache in LDS (0: all the LDS the survivor stacks
 * leave; smaller values trade cache hits for nothing else, since the block's
 * LDS size stays).  For cgbdsm_matching_batch the first query's setting
 * applies. */
cgbdsm_status cgbdsm_set_dfs_steal(cgbdsm_query* q, uint32_t range_min, uint32_t cache_cap);

/* Cooperative deadline for the final DFS of cgbdsm_matching (default 120 s,
 * 0 = none).
 * Past the deadline the host raises a flag the kernel polls; the call then
 * returns CGBDSM_ERR_OVERFLOW because the count is incomplete.  No reference
 * counterpart (the reference waits unboundedly, graph/graph_gpu.cu:1457-1464). */
cgbdsm_status cgbdsm_set_deadline_ms(cgbdsm_query* q, double ms);

/* CPU/GPU split of the final DFS of cgbdsm_matching: the host counts frontier
 * rows [0, k) on `threads` threads while the GPU counts rows [k, n), and the
 * two per-edge counts are added.  Replaces the reference's
 * Step5CPUGPUSplit (graph/graph_gpu.cu:1185-1258) with its host count
 * cpuExtendBFSAllBitCountParallel (index/cpu_dfs.cpp:173-217) and host index
 * mirror (index/cpu_index_mirror.h:32-48, kept in sync at
 * index/calig_helper_gamma.cu:279-302, 503-516, 558-566, 588-595).
 *   threads   0 = off (the default; frees the mirror), < 0 = hardware
 *             threads / NUMA nodes (index/calig.cpp:1092-1095);
 *   ratio     the host's initial share of the rows, in [0, 0.95]; 0 = one
 *             row per host thread on the first split batch;
 *   alpha     weight of the newest measured balance in the moving average
 *             of the share (graph/graph_gpu.cu:1250-1253); 0 keeps the
 *             share fixed;
 *   min_rows  frontiers of fewer rows stay on the GPU (0 = 8192, the
 *             reference's default, graph/graph_gpu.h:128).
 * The share has no lower clamp (the reference's is 0.05): a share under one
 * row leaves the batch to the GPU.  The host mirror of M is built from the
 * device's index at the next batch and follows every batch after it; counts
 * are identical to an unsplit count.  Single-query counts only
 * (cgbdsm_matching, cgbdsm_run_batch); batched launches and tuple output
 * stay on the GPU. */
cgbdsm_status cgbdsm_set_cpu_split(cgbdsm_query* q, int threads, double ratio, double alpha, uint64_t min_rows);
typedef struct {
    int threads;                /* host threads (0: split off)                      */
    double ratio;               /* share the next split batch starts from           */
    uint64_t batches;           /* matchings split so far                           */
    uint64_t last_cpu_rows;     /* rows counted on the host in the last split batch */
    uint64_t last_gpu_rows;
    double last_cpu_ms;         /* host count, wall time                            */
    double last_gpu_ms;         /* GPU share, HIP events on the query's stream      */
    double last_wait_ms;        /* host count's wait for the mirror upkeep          */
    uint64_t last_cpu_matches;  /* matches counted on the host                      */
    uint64_t mirror_entries;    /* entries of M in the host mirror                  */
    double mirror_ms;           /* mirror upkeep of the last batch so far (host
                                   time on its background thread)                  */
} cgbdsm_split_stats;
cgbdsm_status cgbdsm_cpu_split_stats(cgbdsm_query* q, cgbdsm_split_stats* out);
/* The host mirror's M (which = 0) or the batch's U (which = 1) as sorted keys
 * in the boundary format, like cgbdsm_index_export (no reference
 * counterpart: the mirror is internal there). */
cgbdsm_status cgbdsm_cpu_mirror_export(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out);

/* Kernel classes timed with HIP events on the query's stream when profiling
 * is enabled (bytes = algorithmic bytes as defined per class). */
enum {
    CGBDSM_K_SORT = 0,          /* key decode + LSD radix sort              */
    CGBDSM_K_MERGE_INSERT = 1,  /* sorted set-union merge into M lists      */
    CGBDSM_K_MERGE_REMOVE = 2,  /* sorted set-difference compaction of M    */
    CGBDSM_K_UPDATE_INDEX = 3,  /* segment discovery + U lists              */
    CGBDSM_K_MASKS = 4,         /* support masks, all depths                */
    CGBDSM_K_SEED = 5,          /* depth-2 frontier                         */
    CGBDSM_K_EXTEND = 6,        /* BFS levels                               */
    CGBDSM_K_DFS = 7,           /* final DFS count                          */
    CGBDSM_K_GPMA_UPDATE = 8,   /* GPMA layout: route, rebalance, delete    */
    CGBDSM_K_GPMA_EXTRACT = 9,  /* GPMA layout: gap array -> flat lists     */
    CGBDSM_K_COUNT = 10
};
typedef struct {
    double ms;          /* summed device time */
    uint64_t launches;
    uint64_t bytes;     /* algorithmic bytes (merge classes, DFS)          */
    uint64_t items;     /* keys (sort/merge), rows (seed/extend/DFS)        */
} cgbdsm_ktime;
cgbdsm_status cgbdsm_profile(cgbdsm_query* q, int enable);
/* Copies CGBDSM_K_COUNT entries (synchronises the stream); reset != 0 clears. */
cgbdsm_status cgbdsm_kernel_times(cgbdsm_query* q, cgbdsm_ktime* out, int reset);

/* Parity dumps.  which = 0: main index M, 1: update index U.
 * Writes the index as sorted keys in the boundary format into out[cap];
 * *n_out receives the total number of entries (also when cap is too small). */
cgbdsm_status cgbdsm_index_export(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out);

/* cgbdsm_index_export without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Row lengths, scan partials and the exported keys live in one
 * static device array of 32 MiB per device (4 Mi u64 words, held until the
 * process ends): an index whose rows + entries (+ scan partials) exceed it
 * fails with CGBDSM_ERR_LIMIT.  out must hold every entry: cap < *n_out
 * fails with CGBDSM_ERR_INVALID after *n_out is set.  Calls no hipMalloc /
 * hipFree; a GPMA index whose flat lists are stale is refreshed only when
 * that needs no pool growth (else CGBDSM_ERR_LIMIT).  Every call on one
 * device shares the array: calls that overlap on one device (several
 * threads or streams) race and must be serialised by the caller. */
cgbdsm_status cgbdsm_index_export_static(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out);

/* Support-mask words for (edge ei, query vertex qv): ceil(dv/64) u64 words,
 * bit v%64 of word v/64.  *all_ones = 1 for tiles that are not rebuilt. */
cgbdsm_status cgbdsm_mask_export(cgbdsm_query* q, uint32_t ei, uint32_t qv, uint64_t* words,
                                 uint64_t nwords, int* all_ones);

/* Full embeddings of the last cgbdsm_matching call, as rows of qv ids in
 * global-order positions, ei tag stripped (small inputs only: enumerates on
 * the device into out[cap*qv]).  *n_out = number of embeddings. */
cgbdsm_status cgbdsm_match_tuples(cgbdsm_query* q, uint32_t* out, uint64_t cap, uint64_t* n_out,
                                  uint8_t* ei_out);

cgbdsm_status cgbdsm_stats_get(cgbdsm_query* q, cgbdsm_stats* out);

/* Device-side timing of the hot path of one batch already staged
 * (apply + update index + masks + matching + commit), returning per-stage
 * kernel time measured with HIP events on the query's stream. */
cgbdsm_status cgbdsm_run_batch(cgbdsm_query* q, unsigned long long* num_matches,
                               uint64_t* per_edge, int* skipped);

/* Synchronise the query's stream. */
cgbdsm_status cgbdsm_sync(cgbdsm_query* q);

/* ------------------------------------------------------------------------
 * Device Stage 1: the CaLiG producer on the GPU.
 *
 * One data graph per context is shared by every query attached to it (the
 * reference builds one CaLiG, with its own copy of G, per process and query,
 * main_calig.cpp:29-35).  Per batch: cgbdsm_graph_apply_range once, then
 * cgbdsm_calig_update for every attached query, which stages that query's
 * batch exactly as cgbdsm_batch_stage would stage the host producer's keys;
 * the Stage-2 calls follow unchanged.  Lit bits equal the greatest fixpoint
 * of the CaLiG check on the current graph; keys use the same boundary format
 * and compressed ids (ascending original id among newly lit vertices).
 * Deletion streams: a pair of negative ids (-a-1, -b-1) deletes edge (a, b)
 * (CaLiG::updateIndex's deletion branch, index/calig.cpp:980-989, 477-535);
 * lit pairs that lose their support go dark, and the staged Del keys also
 * remove the main-index entries of the deleted edges and of those pairs.
 * Positive matches count the batch's insertions on the graph after the
 * batch; cgbdsm_calig_match_deleted counts the embeddings its deletions
 * destroy.  (The reference's ConstructUpdate indexes vertices with the raw
 * ids, so its GPU stage runs insertion streams only.)  The host producer
 * (cgh_*) skips negative ids.
 * ------------------------------------------------------------------------ */
typedef struct cgbdsm_graph cgbdsm_graph;

/* CaLiG::inputG's G (labels, adjacency; index/calig.cpp:166-216) on the
 * device: n vertices with raw labels[n], m undirected edges as u32 pairs
 * edges[2*m] (self-loops, out-of-range ids and duplicates dropped).
 * n < 2^28. */
cgbdsm_status cgbdsm_graph_create(cgbdsm_ctx* ctx, uint32_t n, const int32_t* labels, const uint32_t* edges,
                                  uint64_t m, cgbdsm_graph** out);
/* Fails while queries are attached (destroy them first). */
cgbdsm_status cgbdsm_graph_destroy(cgbdsm_graph* g);
/* The update stream as CaLiG::inputUpdate holds it (index/calig.cpp:719-733):
 * n_ints ints, two per pair, copied into HBM once. */
cgbdsm_status cgbdsm_graph_set_stream(cgbdsm_graph* g, const int32_t* ints, uint64_t n_ints);
/* G.nei inserts of CaLiG::updateIndex for stream ints [start, start+count)
 * (index/calig.cpp:950-1003; pairs are ints start+2i, start+2i+1 as
 * main_calig.cpp:220-222 slices them, so an odd count shifts the pairing of
 * later batches as in the reference).  Every attached query must have
 * consumed the previous batch. */
cgbdsm_status cgbdsm_graph_apply_range(cgbdsm_graph* g, uint64_t start, uint64_t count);
/* Batches applied, directed adjacency entries, new directed edges of the
 * latest batch.  Insertions of present edges and deletions of absent ones
 * are dropped, both judged against the graph before the batch. */
cgbdsm_status cgbdsm_graph_info(cgbdsm_graph* g, uint64_t* batches, uint64_t* directed_edges, uint64_t* last_new);

/* CaLiG::constructCand + staticFilter (index/calig.cpp:325-356, 705-717)
 * for query q over g, then the initial main index: ConvertGlobalIndex
 * (index/calig_helper_gamma.cu:238-360), as cgbdsm_index_init with the
 * host keys.  qlabels[qv] are the query's raw labels; cardinalities /
 * avg_degrees as cgbdsm_index_init.  Call after cgbdsm_query_create (and
 * cgbdsm_query_set_layout), before cgbdsm_query_set_plan. */
cgbdsm_status cgbdsm_calig_attach(cgbdsm_query* q, cgbdsm_graph* g, const int32_t* qlabels, uint32_t* cardinalities,
                                  float* avg_degrees);
/* Stage 1 of the graph's latest batch for q: CaLiG::updateIndex,
 * ConstructUpdate, mergeDeltas (index/calig.cpp:950-1088) and
 * PrepareUpdateAll / PrepareUpdateIndex (index/calig_helper_gamma.cu:363-445),
 * ending with the batch staged (as cgbdsm_batch_stage). */
cgbdsm_status cgbdsm_calig_update(cgbdsm_query* q);
/* Negative matches of the staged batch (call right after
 * cgbdsm_calig_update, before the Stage-2 calls): the embeddings of the
 * graph before the batch that use one of its deleted edges, with the
 * per-edge split of cgbdsm_matching (the largest query-edge index mapped to
 * a deleted edge).  Runs Stage 2 with the deleted edges' entries as the
 * update index and leaves the positive batch staged; 0 for a batch without
 * deletions.  No reference counterpart (the reference's GPU stage matches
 * insertions only; CaLiG's turn-off path is index/calig.cpp:410-535). */
cgbdsm_status cgbdsm_calig_match_deleted(cgbdsm_query* q, unsigned long long* num_matches, uint64_t* per_edge);
/* Lit bitmask per original vertex (n = the graph's vertex count). */
cgbdsm_status cgbdsm_calig_export_lit(cgbdsm_query* q, uint32_t* li, uint32_t n);
/* Compressed id -> original vertex id; *n_out = compressed vertices. */
cgbdsm_status cgbdsm_calig_old_ids(cgbdsm_query* q, uint32_t* ids, uint32_t cap, uint32_t* n_out);
/* Last cgbdsm_calig_attach / _update: wall ms, and out6 = activation rounds,
 * deletion rounds, pairs lit optimistically, pairs removed, new lit pairs,
 * compressed vertices. */
cgbdsm_status cgbdsm_calig_stats(cgbdsm_query* q, double* ms, uint64_t* out6);
/* Keys of the staged batch in boundary format: which = 0 ins, 1 del,
 * 2 upd (as staged, unsorted); *n_out = count. */
cgbdsm_status cgbdsm_staged_keys(cgbdsm_query* q, int which, uint64_t* out, uint64_t cap, uint64_t* n_out);

#ifdef __cplusplus
}
#endif
#endif /* CGBDSM_H */
