/*
 * cgbdsm_host.h -- C ABI of the host pipeline (Stage-1 CaLiG producer,
 * planner, key encoder, workload generator).
 *
 * The cgcsm_* executables use the same code through C++.  This C surface
 * lets tests and benchmarks drive Stage 1 batch by batch and feed the keys
 * to the device engine (cgbdsm.h) and to the CPU oracle.
 *
 * Reference counterparts: CaLiG (index/calig.h:42-96), Plan
 * (graph/plan.h:38-71), CPUGraphLoader (graph/graph.h:65-83), the host half of
 * CaLiGHelperGamma (index/calig_helper_gamma.h:100-110) and the batch loop of
 * main_calig.cpp:216-275.
 */
#ifndef CGBDSM_HOST_H
#define CGBDSM_HOST_H

#include <stdint.h>

#include "cgbdsm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cgh_session cgh_session;

enum { CGH_QUERY_TREE = 0, CGH_QUERY_DENSE = 1, CGH_QUERY_SPARSE = 2 };

typedef struct {
    uint32_t scale;          /* R-MAT scale: 2^scale vertices              */
    double avg_degree;       /* undirected edges = n * avg_degree / 2       */
    double a, b, c;          /* R-MAT quadrant probabilities (d = 1-a-b-c)  */
    uint32_t num_labels;
    double zipf_s;           /* 0: uniform labels                           */
    double initial_frac;     /* share of edges in the initial graph         */
    uint64_t stream_edges;   /* 0: every remaining edge                     */
    uint64_t seed;
    uint32_t query_size;
    uint32_t query_kind;     /* CGH_QUERY_*                                 */
    uint64_t query_seed;
} cgh_synth_spec;

typedef struct {
    const uint64_t* ins; uint64_t n_ins;
    const uint64_t* del; uint64_t n_del;
    const uint64_t* upd; uint64_t n_upd;
    const uint8_t* edge_ok;  /* [qe]                                        */
    uint32_t dv;             /* compressed vertices after this batch        */
    uint32_t batch;
    uint64_t pairs;          /* stream edges in this batch                  */
    double ms_update_index, ms_construct_update, ms_prepare_all, ms_prepare_index;
} cgh_batch;

const char* cgh_last_error(void);

/* Write query / data-graph / update-stream text files for a synthetic spec. */
int cgh_generate_files(const cgh_synth_spec* spec, const char* query_path, const char* data_path,
                       const char* update_path);
/* The same graph (initial and stream edges together) as a binary CSR prefix
 * (<prefix>.meta.txt, .vertex.bin, .edge.bin) plus the query text file.  A
 * data path that does not end in ".graph" is read as such a prefix
 * (main_calig.cpp:93-101, index/calig.cpp:218-323): labels from rand(),
 * the update stream split off by a hash of each edge. */
int cgh_generate_binary(const cgh_synth_spec* spec, const char* query_path, const char* prefix);

/* Load files (CaLiG::inputQ/inputG/inputUpdate + CPUGraphLoader), then
 * constructCand + staticFilter (main_calig.cpp:88-116).  batch_ints counts
 * ints, two per edge, as the reference CLI does (main_calig.cpp:65-69). */
int cgh_session_open(const char* query_path, const char* data_path, const char* update_path,
                     uint32_t batch_ints, int threads, uint32_t ebits, cgh_session** out);
/* Same from an in-memory synthetic spec (no files). */
int cgh_session_open_synth(const cgh_synth_spec* spec, uint32_t batch_ints, int threads, uint32_t ebits,
                           cgh_session** out);
void cgh_session_close(cgh_session* s);

/* Query model + per-edge indexing orders (before the matching order). */
int cgh_query_tables(cgh_session* s, cgbdsm_plan* out);
/* Initial keys of every lit candidate entry (ConvertGlobalIndex host half);
 * the pointer stays valid until the session closes. */
int cgh_initial_keys(cgh_session* s, const uint64_t** keys, uint64_t* n, uint32_t* dv, uint32_t* card,
                     float* avg);
/* GenerateMatchingOrders_v2 from cardinalities / average degrees. */
int cgh_make_plan(cgh_session* s, const uint32_t* card, const float* avg, cgbdsm_plan* out);

uint32_t cgh_num_batches(cgh_session* s);
uint32_t cgh_num_threads(cgh_session* s);
/* Run Stage 1 for the next batch; pointers stay valid until the next call. */
int cgh_next_batch(cgh_session* s, cgh_batch* out);
/* Stage 1 for a batch of the next n_edges stream edges (fewer at the end of
 * the stream): batches of any size, e.g. a serving loop's arrivals.  The
 * stream is consumed in order by both calls. */
int cgh_next_batch_edges(cgh_session* s, uint64_t n_edges, cgh_batch* out);
/* Stream ints (two per edge) not yet consumed by Stage 1. */
uint64_t cgh_stream_left(cgh_session* s);
/* The whole update stream as loaded (ints, two per edge, in stream order);
 * *n_out = its length, up to cap ints copied. */
int cgh_export_updates(cgh_session* s, uint32_t* out, uint64_t cap, uint64_t* n_out);

/* Compressed id -> original vertex id. */
int cgh_old_ids(cgh_session* s, const uint32_t** ids, uint32_t* n);
/* Current data graph seen by CaLiG (label-filtered, CSR; offs has n+1). */
int cgh_export_graph(cgh_session* s, uint64_t* offs, uint32_t* adj, uint64_t adj_cap, int32_t* labels,
                     uint32_t* n_out, uint64_t* m_out);
/* Lit bitmask per original vertex. */
int cgh_export_li(cgh_session* s, uint32_t* li, uint32_t n);
/* cand[ui][uj] of v as a sorted list (original ids); *n_out = size or -1. */
int cgh_export_cand(cgh_session* s, uint32_t v, uint32_t ui, uint32_t uj, uint32_t* out, uint32_t cap,
                    int64_t* n_out);
/* CaLiG's query: raw labels[qv], edges as (u,uu) pairs in qe_list order. */
int cgh_query_info(cgh_session* s, int32_t* labels, uint32_t* edges_uv, uint32_t* qv, uint32_t* qe);

/* ------------------------------------------------------------------------
 * Workloads and planners for the device Stage 1 (cgbdsm_calig_*), which
 * needs no host CaLiG: the synthetic graph and stream of a spec (the same
 * graph cgh_session_open_synth builds), random-walk queries sampled on it,
 * and the planner of a query given by labels and edges.
 * ------------------------------------------------------------------------ */
typedef struct cgh_workload cgh_workload;
typedef struct cgh_planner cgh_planner;

int cgh_workload_open(const cgh_synth_spec* spec, cgh_workload** out);
void cgh_workload_close(cgh_workload* w);
/* n vertices, raw labels[n], initial edges as m u32 pairs; pointers stay
 * valid until the workload closes. */
int cgh_workload_graph(cgh_workload* w, uint32_t* n, const int32_t** labels, const uint32_t** edges, uint64_t* m);
/* The update stream as ints, two per pair, in stream order. */
int cgh_workload_stream(cgh_workload* w, const int32_t** ints, uint64_t* n_ints);
/* A query of `size` vertices and kind CGH_QUERY_* sampled by a seeded random
 * walk on the whole graph (initial and stream edges): raw labels[qv] and
 * edges as (u, uu) pairs in qe_list order. */
int cgh_workload_query(cgh_workload* w, uint32_t size, uint32_t kind, uint64_t seed, int32_t* labels,
                       uint32_t* edges_uv, uint32_t* qv, uint32_t* qe);

/* Planner (Plan, graph/plan.h:38-71) of a query given by labels and edges. */
int cgh_planner_open(const int32_t* labels, uint32_t qv, const uint32_t* edges_uv, uint32_t qe, uint32_t ebits,
                     cgh_planner** out);
void cgh_planner_close(cgh_planner* p);
/* As cgh_query_tables / cgh_make_plan. */
int cgh_planner_tables(cgh_planner* p, cgbdsm_plan* out);
int cgh_planner_make_plan(cgh_planner* p, const uint32_t* card, const float* avg, cgbdsm_plan* out);

#ifdef __cplusplus
}
#endif
#endif /* CGBDSM_HOST_H */
