/*
 * Oracle masks and merged matching -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Masks: kernels/indexing.cu:2295-2379 as a definition.  For edge ei with
 * indexing order io (io.vs[0], io.vs[1] = ei's endpoints):
 *   depth 0/1:  bit v of (ei, vs0) <=> U[vs0->vs1][v] non-empty (same for vs1)
 *   depth d>=2, rebuilt vertex u0: bit v <=> for every backward neighbour u1,
 *     some nb in M[u0->u1][v] (or in U[u0->u1][v] when that edge's index is
 *     <= ei) has bit nb set in (ei, u1)
 *   tiles of vertices that are not rebuilt are all ones.
 *
 * Matching: seeding kernels/enumeration.cu:727-809, extension
 * index/cpu_dfs.cpp:81-132 plus the support-mask prune.  List of directed
 * edge `dir` at owner v under tag ei: base = U if dir is one of ei's two
 * directions, else M; appended U when edge(dir) < ei.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_internal.h"

#define QV 12

uint64_t or_mask_words(const or_index* x) { return ((uint64_t)x->dv + 63) / 64; }

static inline int tile_bit(const uint64_t* masks, const uint8_t* all_ones, uint64_t words, uint32_t ei, uint32_t qv,
                           uint32_t v) {
    if (!masks) return 1;
    size_t t = (size_t)ei * QV + qv;
    if (all_ones[t]) return 1;
    return (int)((masks[t * words + (v >> 6)] >> (v & 63)) & 1u);
}

int or_masks(const or_index* x, const cgbdsm_plan* p, uint64_t* out, uint8_t* all_ones) {
    const uint64_t W = or_mask_words(x);
    memset(out, 0, sizeof(uint64_t) * W * p->qe * QV);
    memset(all_ones, 0, (size_t)p->qe * QV);
    for (uint32_t ei = 0; ei < p->qe; ei++) {
        for (uint32_t qv = 0; qv < p->qv; qv++)
            if (!p->rebuild_v[ei][qv]) all_ones[ei * QV + qv] = 1;
        const uint8_t* vs = p->io_vs[ei];
        uint32_t u0 = vs[0], u1 = vs[1];
        uint32_t i01 = p->eidx[u0 * QV + u1], i10 = p->eidx[u1 * QV + u0];
        for (uint32_t v = 0; v < x->dv; v++) {
            uint32_t l;
            or_upd_list(x, i01, v, &l);
            if (l) out[((size_t)ei * QV + u0) * W + (v >> 6)] |= 1ull << (v & 63);
            or_upd_list(x, i10, v, &l);
            if (l) out[((size_t)ei * QV + u1) * W + (v >> 6)] |= 1ull << (v & 63);
        }
        for (uint32_t d = 2; d < p->qv; d++) {
            uint32_t a = vs[d];
            if (!p->rebuild_v[ei][a]) continue;
            uint64_t* tile = out + ((size_t)ei * QV + a) * W;
            for (uint32_t v = 0; v < x->dv; v++) {
                int ok = 1;
                for (uint32_t k = p->io_bni_offs[ei][d]; k < p->io_bni_offs[ei][d + 1] && ok; k++) {
                    uint32_t b = vs[p->io_bni[ei][k]];
                    uint32_t bwd = p->eidx[a * QV + b];
                    int sat = 0;
                    uint32_t l;
                    const uint32_t* L = or_main_list(x, bwd, v, &l);
                    for (uint32_t j = 0; j < l && !sat; j++) sat = tile_bit(out, all_ones, W, ei, b, L[j]);
                    if (!sat && p->dir_to_edge[bwd] <= ei) {
                        L = or_upd_list(x, bwd, v, &l);
                        for (uint32_t j = 0; j < l && !sat; j++) sat = tile_bit(out, all_ones, W, ei, b, L[j]);
                    }
                    ok = sat;
                }
                if (ok) tile[v >> 6] |= 1ull << (v & 63);
            }
        }
    }
    return 0;
}

typedef struct {
    const or_index* x;
    const cgbdsm_plan* p;
    const uint64_t* masks;
    const uint8_t* all_ones;
    uint64_t W;
    uint32_t ei, sa, sb;
    uint32_t matched[QV];
    uint64_t cnt;
    uint32_t* tuples;
    uint8_t* tuple_ei;
    uint64_t cap, ntup;
} runner;

static inline void resolve(const runner* r, uint32_t dir, uint32_t v, const uint32_t** base, uint32_t* bl,
                           const uint32_t** app, uint32_t* al) {
    if (dir == r->sa || dir == r->sb) *base = or_upd_list(r->x, dir, v, bl);
    else *base = or_main_list(r->x, dir, v, bl);
    *app = NULL;
    *al = 0;
    if (r->p->dir_to_edge[dir] < r->ei) *app = or_upd_list(r->x, dir, v, al);
}

static inline int sorted_has(const uint32_t* a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t m = (lo + hi) / 2;
        if (a[m] < v) lo = m + 1;
        else hi = m;
    }
    return lo < n && a[lo] == v;
}

static void emit(runner* r) {
    if (r->tuples && r->ntup < r->cap) {
        memcpy(r->tuples + r->ntup * r->p->qv, r->matched, sizeof(uint32_t) * r->p->qv);
        r->tuple_ei[r->ntup] = (uint8_t)r->ei;
    }
    r->ntup++;
}

static void dfs(runner* r, uint32_t d) {
    const cgbdsm_plan* p = r->p;
    uint32_t ud = p->go_vs[d];
    uint32_t s = p->go_bni_offs[d], e = p->go_bni_offs[d + 1];
    uint32_t pq = p->go_bni[s];
    uint32_t pe = p->eidx[p->go_vs[pq] * QV + ud];
    const uint32_t *base, *app;
    uint32_t bl, al;
    resolve(r, pe, r->matched[pq], &base, &bl, &app, &al);
    const uint32_t* ob[QV];
    const uint32_t* oa[QV];
    uint32_t obl[QV], oal[QV];
    uint32_t no = 0;
    for (uint32_t k = s + 1; k < e; k++, no++) {
        uint32_t bq = p->go_bni[k];
        resolve(r, p->eidx[p->go_vs[bq] * QV + ud], r->matched[bq], &ob[no], &obl[no], &oa[no], &oal[no]);
    }
    for (uint32_t t = 0; t < bl + al; t++) {
        uint32_t nbr = t < bl ? base[t] : app[t - bl];
        if (!tile_bit(r->masks, r->all_ones, r->W, r->ei, ud, nbr)) continue;
        int ok = 1;
        for (uint32_t i = 0; i < d && ok; i++)
            if (r->matched[i] == nbr) ok = 0;
        for (uint32_t k = 0; k < no && ok; k++)
            ok = sorted_has(ob[k], obl[k], nbr) || sorted_has(oa[k], oal[k], nbr);
        if (!ok) continue;
        r->matched[d] = nbr;
        if (d + 1 == p->qv) {
            r->cnt++;
            emit(r);
        } else {
            dfs(r, d + 1);
        }
    }
}

/* Seeds of tag ei for owners [v_lo, v_hi) and their DFS. */
static void seed_range(runner* r, uint32_t v_lo, uint32_t v_hi) {
    const cgbdsm_plan* p = r->p;
    uint32_t q0 = p->go_vs[0], q1 = p->go_vs[1];
    uint32_t idx = p->eidx[q0 * QV + q1];
    for (uint32_t v = v_lo; v < v_hi; v++) {
        if (!tile_bit(r->masks, r->all_ones, r->W, r->ei, q0, v)) continue;
        const uint32_t *base, *app;
        uint32_t bl, al;
        resolve(r, idx, v, &base, &bl, &app, &al);
        for (uint32_t t = 0; t < bl + al; t++) {
            uint32_t nbr = t < bl ? base[t] : app[t - bl];
            if (!tile_bit(r->masks, r->all_ones, r->W, r->ei, q1, nbr)) continue;
            r->matched[0] = v;
            r->matched[1] = nbr;
            if (p->qv == 2) {
                r->cnt++;
                emit(r);
            } else {
                dfs(r, 2);
            }
        }
    }
}

typedef struct {
    runner proto;
    const uint8_t* edge_ok;
    uint32_t dv, qe;
    uint64_t chunk;
    volatile uint64_t next;
    pthread_mutex_t mu;
    uint64_t per_edge[CGBDSM_MAX_QE];
} pool_t;

static void* worker(void* arg) {
    pool_t* P = (pool_t*)arg;
    uint64_t local[CGBDSM_MAX_QE];
    memset(local, 0, sizeof(local));
    const uint64_t per_ei = (P->dv + P->chunk - 1) / P->chunk;
    const uint64_t total = per_ei * P->qe;
    for (;;) {
        uint64_t t = __atomic_fetch_add(&P->next, 1, __ATOMIC_RELAXED);
        if (t >= total) break;
        uint32_t ei = (uint32_t)(t / per_ei);
        if (!P->edge_ok[ei]) continue;
        uint64_t lo = (t % per_ei) * P->chunk, hi = lo + P->chunk;
        if (hi > P->dv) hi = P->dv;
        runner r = P->proto;
        r.ei = ei;
        r.sa = P->proto.p->eidx[P->proto.p->qe_u[ei] * QV + P->proto.p->qe_v[ei]];
        r.sb = P->proto.p->eidx[P->proto.p->qe_v[ei] * QV + P->proto.p->qe_u[ei]];
        r.cnt = 0;
        seed_range(&r, (uint32_t)lo, (uint32_t)hi);
        local[ei] += r.cnt;
    }
    pthread_mutex_lock(&P->mu);
    for (uint32_t e = 0; e < P->qe; e++) P->per_edge[e] += local[e];
    pthread_mutex_unlock(&P->mu);
    return NULL;
}

uint64_t or_match(const or_index* x, const cgbdsm_plan* p, const uint64_t* masks, const uint8_t* all_ones,
                  const uint8_t* edge_ok, uint64_t* per_edge, uint32_t* tuples, uint8_t* tuple_ei, uint64_t cap,
                  int threads) {
    runner proto;
    memset(&proto, 0, sizeof(proto));
    proto.x = x;
    proto.p = p;
    proto.masks = masks;
    proto.all_ones = all_ones;
    proto.W = or_mask_words(x);
    uint64_t total = 0;
    if (tuples || threads <= 1) {
        proto.tuples = tuples;
        proto.tuple_ei = tuple_ei;
        proto.cap = cap;
        for (uint32_t ei = 0; ei < p->qe; ei++) {
            uint64_t c = 0;
            if (edge_ok[ei]) {
                proto.ei = ei;
                proto.sa = p->eidx[p->qe_u[ei] * QV + p->qe_v[ei]];
                proto.sb = p->eidx[p->qe_v[ei] * QV + p->qe_u[ei]];
                proto.cnt = 0;
                seed_range(&proto, 0, x->dv);
                c = proto.cnt;
            }
            if (per_edge) per_edge[ei] = c;
            total += c;
        }
        return total;
    }
    pool_t P;
    memset(&P, 0, sizeof(P));
    P.proto = proto;
    P.edge_ok = edge_ok;
    P.dv = x->dv;
    P.qe = p->qe;
    P.chunk = 256;
    pthread_mutex_init(&P.mu, NULL);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, worker, &P);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&P.mu);
    for (uint32_t e = 0; e < p->qe; e++) {
        if (per_edge) per_edge[e] = P.per_edge[e];
        total += P.per_edge[e];
    }
    return total;
}
