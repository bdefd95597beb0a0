/*
 * CPU oracle for the Stage-2 hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or the timed CPU baseline).
 * The product path (cgbdsm_amd) never calls it.
 *
 * It restates the reference CPU path in plain C:
 *   index algebra   index/cpu_index_mirror.cpp:132-269 (set union on insert,
 *                   set difference on remove, unsorted-unique update lists)
 *                   and index/calig_helper_gamma.cu:485-567 (call order)
 *   support masks   kernels/indexing.cu:2295-2379 (definition, all tiles)
 *   seeding         kernels/enumeration.cu:727-809
 *   final DFS       index/cpu_dfs.cpp:67-132 (+ the mask prune of
 *                   kernels/enumeration.cu:928-935)
 * and adds independent checkers that do not use the candidate index:
 *   brute force     number of embeddings of Q in G (injective, label and
 *                   edge preserving, non-induced)
 *   anchored delta  per query edge e, embeddings of G_b whose largest query
 *                   edge mapped to a batch edge is e
 *   LI fixpoint     greatest fixpoint of CaLiG's local candidate condition
 *                   (index/calig.cpp:385-408 on a from-scratch index)
 *
 * Parity pins (see tests/): closed form emb(G_b)-emb(G_{b-1}) == sum of
 * anchored counts == oracle match count; textbook cases (triangles, paths,
 * stars); set algebra against sort/unique; masks against brute recomputation.
 */
#ifndef CGBDSM_ORACLE_H
#define CGBDSM_ORACLE_H

#include <stdint.h>

#include "../include/cgbdsm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_index or_index;

or_index* or_index_new(uint32_t ndir, uint32_t ebits);
void or_index_free(or_index* x);
void or_index_set_dv(or_index* x, uint32_t dv);
uint32_t or_index_dv(const or_index* x);
/* M <- set(keys) */
int or_index_init(or_index* x, const uint64_t* keys, uint64_t n);
/* M <- M u set(keys) */
int or_index_insert(or_index* x, const uint64_t* keys, uint64_t n);
/* M <- M \ keys */
int or_index_remove(or_index* x, const uint64_t* keys, uint64_t n);
/* U <- keys, sorted, duplicates kept */
int or_index_rebuild_update(or_index* x, const uint64_t* keys, uint64_t n);
/* which 0 = M, 1 = U; sorted boundary keys; returns total entries. */
uint64_t or_index_export(const or_index* x, int which, uint64_t* out, uint64_t cap);
/* touched-row statistics of the last insert/remove (for byte accounting) */
uint64_t or_index_entries(const or_index* x, int which);

/* Support masks for every (ei, qv): out[(ei*12+qv)*words + w], words =
 * ceil(dv/64); all_ones[ei*12+qv] = 1 for tiles that are not rebuilt. */
int or_masks(const or_index* x, const cgbdsm_plan* p, uint64_t* out, uint8_t* all_ones);
uint64_t or_mask_words(const or_index* x);

/* Merged matching count (seed + DFS) for one batch.  masks may be NULL (no
 * prune).  per_edge[qe] receives counts per tag.  tuples (optional) receives
 * up to cap embeddings as qv compressed ids in global-order positions,
 * tuple_ei[i] the tag.  Returns the total count.  threads > 1 splits the
 * seed rows over pthreads. */
uint64_t or_match(const or_index* x, const cgbdsm_plan* p, const uint64_t* masks, const uint8_t* all_ones,
                  const uint8_t* edge_ok, uint64_t* per_edge, uint32_t* tuples, uint8_t* tuple_ei, uint64_t cap,
                  int threads);

/* Number of embeddings of the query in a CSR graph (adj sorted per vertex). */
uint64_t or_brute_count(uint32_t n, const uint64_t* offs, const uint32_t* adj, const int32_t* labels, uint32_t qv,
                        const int32_t* qlabels, uint32_t qe, const uint32_t* qedges);

/* Anchored per-edge counts of new embeddings for one batch.  The graph is
 * G_b (after the batch); is_new(a,b) marks the batch's edges, given as a
 * sorted list of (min<<32|max) keys.  qedges are in qe_list order.
 * per_edge[qe] receives the counts; returns the total. */
uint64_t or_anchored_count(uint32_t n, const uint64_t* offs, const uint32_t* adj, const int32_t* labels,
                           uint32_t qv, const int32_t* qlabels, uint32_t qe, const uint32_t* qedges,
                           const uint64_t* batch_keys, uint64_t n_batch, uint64_t* per_edge, int threads);

/* The new embeddings of one batch themselves, found as or_anchored_count
 * finds them (single-threaded): up to cap rows of qv original vertex ids in
 * query-vertex order go to out[], the tag (largest query edge, qe_list
 * order, mapped to a batch edge) of each to tags[].  Returns the number of
 * new embeddings (also when it exceeds cap). */
uint64_t or_brute_tuples(uint32_t n, const uint64_t* offs, const uint32_t* adj, const int32_t* labels, uint32_t qv,
                         const int32_t* qlabels, uint32_t qe, const uint32_t* qedges, const uint64_t* batch_keys,
                         uint64_t n_batch, uint32_t* out, uint8_t* tags, uint64_t cap);

/* Greatest fixpoint of the lit condition on G: li_out[v] bit u set iff v
 * has u's label and, for every query neighbour uj of u, v has a neighbour
 * lit for uj, with distinct supports for query neighbours sharing a label. */
int or_li_fixpoint(uint32_t n, const uint64_t* offs, const uint32_t* adj, const int32_t* labels, uint32_t qv,
                   const int32_t* qlabels, uint32_t qe, const uint32_t* qedges, uint32_t* li_out);

#ifdef __cplusplus
}
#endif
#endif
