/*
 * Index-free checkers -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * or_brute_count      embeddings of Q in G by plain backtracking: injective,
 *                     label preserving, every query edge maps to a data edge.
 * or_anchored_count   new embeddings of one batch split by the edge tag: the
 *                     largest query edge (qe_list order) that maps to a batch
 *                     edge.  Query edges above the tag must map to old edges,
 *                     the tag edge to a batch edge, lower edges to any edge --
 *                     the counting rule stated by support_material.pdf
 *                     Algorithm S9 and implemented by
 *                     kernels/enumeration.cu:768-791, 937, 1082.
 * or_li_fixpoint      greatest fixpoint of the CaLiG lit condition.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    uint32_t n;
    const uint64_t* offs;
    const uint32_t* adj;
    const int32_t* labels;
    uint32_t qv, qe;
    const int32_t* ql;
    const uint32_t* qedges;
    uint32_t order[CGBDSM_MAX_QV];   /* matching order (connected)            */
    uint32_t pos[CGBDSM_MAX_QV];     /* query vertex -> position              */
    /* for position k: earlier-positioned neighbours and the edge ids */
    uint32_t nb[CGBDSM_MAX_QV][CGBDSM_MAX_QV], nbe[CGBDSM_MAX_QV][CGBDSM_MAX_QV], nnb[CGBDSM_MAX_QV];
    const uint64_t* batch;
    uint64_t n_batch;
    uint32_t tag;                    /* anchored mode: tag edge, else UINT32_MAX */
    uint32_t img[CGBDSM_MAX_QV];
    /* recording (or_brute_tuples): rows of qv ids in query-vertex order */
    uint32_t* rec;
    uint8_t* rec_tag;
    uint64_t rec_cap, rec_n;
} bt_t;

static int has_edge(const bt_t* b, uint32_t x, uint32_t y) {
    uint64_t lo = b->offs[x], hi = b->offs[x + 1];
    while (lo < hi) {
        uint64_t m = (lo + hi) / 2;
        if (b->adj[m] < y) lo = m + 1;
        else hi = m;
    }
    return lo < b->offs[x + 1] && b->adj[lo] == y;
}

static int is_new(const bt_t* b, uint32_t x, uint32_t y) {
    uint64_t k = x < y ? ((uint64_t)x << 32) | y : ((uint64_t)y << 32) | x;
    uint64_t lo = 0, hi = b->n_batch;
    while (lo < hi) {
        uint64_t m = (lo + hi) / 2;
        if (b->batch[m] < k) lo = m + 1;
        else hi = m;
    }
    return lo < b->n_batch && b->batch[lo] == k;
}

/* Edge constraint in anchored mode. */
static int edge_ok(const bt_t* b, uint32_t e, uint32_t x, uint32_t y) {
    if (!has_edge(b, x, y)) return 0;
    if (b->tag == UINT32_MAX) return 1;
    int nw = is_new(b, x, y);
    if (e > b->tag) return !nw;
    if (e == b->tag) return nw;
    return 1;
}

static void setup_order(bt_t* b, uint32_t first, uint32_t second) {
    uint32_t seen[CGBDSM_MAX_QV] = {0};
    uint32_t k = 0;
    b->order[k++] = first;
    seen[first] = 1;
    if (second != UINT32_MAX) {
        b->order[k++] = second;
        seen[second] = 1;
    }
    /* BFS order over the query from the seeds */
    for (uint32_t h = 0; h < k && k < b->qv; h++) {
        uint32_t x = b->order[h];
        for (uint32_t e = 0; e < b->qe; e++) {
            uint32_t a = b->qedges[2 * e], c = b->qedges[2 * e + 1];
            uint32_t y = a == x ? c : (c == x ? a : UINT32_MAX);
            if (y != UINT32_MAX && !seen[y]) {
                seen[y] = 1;
                b->order[k++] = y;
            }
        }
    }
    for (uint32_t i = 0; i < b->qv; i++) b->pos[b->order[i]] = i;
    for (uint32_t i = 0; i < b->qv; i++) {
        b->nnb[i] = 0;
        uint32_t u = b->order[i];
        for (uint32_t e = 0; e < b->qe; e++) {
            uint32_t a = b->qedges[2 * e], c = b->qedges[2 * e + 1];
            uint32_t y = a == u ? c : (c == u ? a : UINT32_MAX);
            if (y != UINT32_MAX && b->pos[y] < i) {
                b->nb[i][b->nnb[i]] = y;
                b->nbe[i][b->nnb[i]] = e;
                b->nnb[i]++;
            }
        }
    }
}

static uint64_t extend(bt_t* b, uint32_t k) {
    if (k == b->qv) {
        if (b->rec) {
            if (b->rec_n < b->rec_cap) {
                memcpy(b->rec + b->rec_n * b->qv, b->img, sizeof(uint32_t) * b->qv);
                b->rec_tag[b->rec_n] = (uint8_t)b->tag;
            }
            b->rec_n++;
        }
        return 1;
    }
    uint32_t u = b->order[k];
    uint64_t cnt = 0;
    if (b->nnb[k] == 0) {
        for (uint32_t v = 0; v < b->n; v++) {
            if (b->labels[v] != b->ql[u]) continue;
            int ok = 1;
            for (uint32_t i = 0; i < k && ok; i++)
                if (b->img[b->order[i]] == v) ok = 0;
            if (!ok) continue;
            b->img[u] = v;
            cnt += extend(b, k + 1);
        }
        return cnt;
    }
    uint32_t p = b->nb[k][0];
    uint32_t pv = b->img[p];
    for (uint64_t t = b->offs[pv]; t < b->offs[pv + 1]; t++) {
        uint32_t v = b->adj[t];
        if (b->labels[v] != b->ql[u]) continue;
        int ok = 1;
        for (uint32_t i = 0; i < k && ok; i++)
            if (b->img[b->order[i]] == v) ok = 0;
        for (uint32_t j = 0; j < b->nnb[k] && ok; j++) ok = edge_ok(b, b->nbe[k][j], b->img[b->nb[k][j]], v);
        if (!ok) continue;
        b->img[u] = v;
        cnt += extend(b, k + 1);
    }
    return cnt;
}

uint64_t or_brute_count(uint32_t n, const uint64_t* offs, const uint32_t* adj, const int32_t* labels, uint32_t qv,
                        const int32_t* qlabels, uint32_t qe, const uint32_t* qedges) {
    bt_t b;
    memset(&b, 0, sizeof(b));
    b.n = n;
    b.offs = offs;
    b.adj = adj;
    b.labels = labels;
    b.qv = qv;
    b.qe = qe;
    b.ql = qlabels;
    b.qedges = qedges;
    b.tag = UINT32_MAX;
    setup_order(&b, 0, UINT32_MAX);
    return extend(&b, 0);
}

typedef struct {
    bt_t proto;
    volatile uint64_t next;
    uint64_t per_edge[CGBDSM_MAX_QE];
    pthread_mutex_t mu;
} anch_pool;

static uint64_t anchored_one(bt_t* b, uint32_t e, uint64_t bi) {
    uint32_t x = (uint32_t)(b->batch[bi] >> 32), y = (uint32_t)b->batch[bi];
    uint32_t qa = b->qedges[2 * e], qb = b->qedges[2 * e + 1];
    uint64_t c = 0;
    for (int o = 0; o < 2; o++) {
        uint32_t fa = o ? y : x, fb = o ? x : y;
        if (b->labels[fa] != b->ql[qa] || b->labels[fb] != b->ql[qb]) continue;
        if (!has_edge(b, fa, fb)) continue;
        b->img[qa] = fa;
        b->img[qb] = fb;
        c += extend(b, 2);
    }
    return c;
}

static void* anch_worker(void* arg) {
    anch_pool* P = (anch_pool*)arg;
    uint64_t local[CGBDSM_MAX_QE];
    memset(local, 0, sizeof(local));
    const uint64_t total = P->proto.n_batch * P->proto.qe;
    for (;;) {
        uint64_t t = __atomic_fetch_add(&P->next, 64, __ATOMIC_RELAXED);
        if (t >= total) break;
        uint64_t te = t + 64 < total ? t + 64 : total;
        for (uint64_t w = t; w < te; w++) {
            uint32_t e = (uint32_t)(w / P->proto.n_batch);
            uint64_t bi = w % P->proto.n_batch;
            bt_t b = P->proto;
            b.tag = e;
            setup_order(&b, b.qedges[2 * e], b.qedges[2 * e + 1]);
            local[e] += anchored_one(&b, e, bi);
        }
    }
    pthread_mutex_lock(&P->mu);
    for (uint32_t e = 0; e < P->proto.qe; e++) P->per_edge[e] += local[e];
    pthread_mutex_unlock(&P->mu);
    return NULL;
}

uint64_t or_anchored_count(uint32_t n, const uint64_t* offs, const uint32_t* adj, const int32_t* labels,
                           uint32_t qv, const int32_t* qlabels, uint32_t qe, const uint32_t* qedges,
                           const uint64_t* batch_keys, uint64_t n_batch, uint64_t* per_edge, int threads) {
    anch_pool P;
    memset(&P, 0, sizeof(P));
    bt_t* b = &P.proto;
    b->n = n;
    b->offs = offs;
    b->adj = adj;
    b->labels = labels;
    b->qv = qv;
    b->qe = qe;
    b->ql = qlabels;
    b->qedges = qedges;
    b->batch = batch_keys;
    b->n_batch = n_batch;
    if (threads < 1) threads = 1;
    pthread_mutex_init(&P.mu, NULL);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, anch_worker, &P);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    pthread_mutex_destroy(&P.mu);
    uint64_t tot = 0;
    for (uint32_t e = 0; e < qe; e++) {
        if (per_edge) per_edge[e] = P.per_edge[e];
        tot += P.per_edge[e];
    }
    return tot;
}

uint64_t or_brute_tuples(uint32_t n, const uint64_t* offs, const uint32_t* adj, const int32_t* labels, uint32_t qv,
                         const int32_t* qlabels, uint32_t qe, const uint32_t* qedges, const uint64_t* batch_keys,
                         uint64_t n_batch, uint32_t* out, uint8_t* tags, uint64_t cap) {
    bt_t b;
    memset(&b, 0, sizeof(b));
    b.n = n;
    b.offs = offs;
    b.adj = adj;
    b.labels = labels;
    b.qv = qv;
    b.qe = qe;
    b.ql = qlabels;
    b.qedges = qedges;
    b.batch = batch_keys;
    b.n_batch = n_batch;
    b.rec = out;
    b.rec_tag = tags;
    b.rec_cap = out && tags ? cap : 0;
    uint64_t total = 0;
    for (uint32_t e = 0; e < qe; e++) {
        b.tag = e;
        setup_order(&b, qedges[2 * e], qedges[2 * e + 1]);
        for (uint64_t bi = 0; bi < n_batch; bi++) total += anchored_one(&b, e, bi);
    }
    return total;
}

/* ---------------------------------------------------------------- LI */

/* Distinct supports for a group of query neighbours sharing a label:
 * backtracking assignment, the tryNei rule of index/calig.cpp:358-383. */
static int assign(const uint32_t* const* sup, const uint32_t* nsup, uint32_t g, uint32_t k, uint32_t* used) {
    if (k == g) return 1;
    for (uint32_t i = 0; i < nsup[k]; i++) {
        uint32_t v = sup[k][i];
        int dup = 0;
        for (uint32_t j = 0; j < k; j++)
            if (used[j] == v) dup = 1;
        if (dup) continue;
        used[k] = v;
        if (assign(sup, nsup, g, k + 1, used)) return 1;
    }
    return 0;
}

int or_li_fixpoint(uint32_t n, const uint64_t* offs, const uint32_t* adj, const int32_t* labels, uint32_t qv,
                   const int32_t* qlabels, uint32_t qe, const uint32_t* qedges, uint32_t* li) {
    uint32_t qn[CGBDSM_MAX_QV][CGBDSM_MAX_QV], nq[CGBDSM_MAX_QV];
    memset(nq, 0, sizeof(nq));
    for (uint32_t e = 0; e < qe; e++) {
        uint32_t a = qedges[2 * e], c = qedges[2 * e + 1];
        qn[a][nq[a]++] = c;
        qn[c][nq[c]++] = a;
    }
    for (uint32_t v = 0; v < n; v++) {
        li[v] = 0;
        for (uint32_t u = 0; u < qv; u++)
            if (labels[v] >= 0 && labels[v] == qlabels[u]) li[v] |= 1u << u;
    }
    uint32_t maxdeg = 0;
    for (uint32_t v = 0; v < n; v++)
        if (offs[v + 1] - offs[v] > maxdeg) maxdeg = (uint32_t)(offs[v + 1] - offs[v]);
    uint32_t* buf = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(maxdeg + 1) * CGBDSM_MAX_QV);
    int changed = 1;
    while (changed) {
        changed = 0;
        for (uint32_t v = 0; v < n; v++) {
            for (uint32_t u = 0; u < qv; u++) {
                if (!(li[v] & (1u << u))) continue;
                const uint32_t* sup[CGBDSM_MAX_QV];
                uint32_t ns[CGBDSM_MAX_QV];
                int ok = 1;
                for (uint32_t k = 0; k < nq[u] && ok; k++) {
                    uint32_t uj = qn[u][k];
                    uint32_t* s = buf + (size_t)k * (maxdeg + 1);
                    uint32_t c = 0;
                    for (uint64_t t = offs[v]; t < offs[v + 1]; t++) {
                        uint32_t w = adj[t];
                        if (labels[w] == qlabels[uj] && (li[w] & (1u << uj))) s[c++] = w;
                    }
                    sup[k] = s;
                    ns[k] = c;
                    if (!c) ok = 0;
                }
                /* same-label groups need distinct supports */
                for (uint32_t k = 0; k < nq[u] && ok; k++) {
                    int32_t lb = qlabels[qn[u][k]];
                    int first = 1;
                    for (uint32_t j = 0; j < k; j++)
                        if (qlabels[qn[u][j]] == lb) first = 0;
                    if (!first) continue;
                    const uint32_t* gs[CGBDSM_MAX_QV];
                    uint32_t gn[CGBDSM_MAX_QV], g = 0, used[CGBDSM_MAX_QV];
                    for (uint32_t j = 0; j < nq[u]; j++)
                        if (qlabels[qn[u][j]] == lb) {
                            gs[g] = sup[j];
                            gn[g] = ns[j];
                            g++;
                        }
                    if (g > 1 && !assign(gs, gn, g, 0, used)) ok = 0;
                }
                if (!ok) {
                    li[v] &= ~(1u << u);
                    changed = 1;
                }
            }
        }
    }
    free(buf);
    return 0;
}
