/*
 * Oracle index algebra -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates CPUIndexMirror (index/cpu_index_mirror.cpp):
 *   MergeInsert   :155-205  sort + unique + std::set_union per (dir, owner)
 *   MergeRemove   :207-229  sort (duplicates kept) + std::set_difference
 *   RebuildUpdate :231-269  sparse per-direction lists, duplicates kept
 * Key layout: (dir << (64-ebits)) | (owner << 32) | target
 * (index/calig_helper_gamma.cu:264); numeric order is (dir, owner, target).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_internal.h"

static void row_reserve(or_row* r, uint32_t need) {
    if (need <= r->cap) return;
    uint32_t c = r->cap ? r->cap : 4;
    while (c < need) c *= 2;
    r->d = (uint32_t*)realloc(r->d, (size_t)c * sizeof(uint32_t));
    r->cap = c;
}

/* LSD radix sort of u64 keys, 16-bit digits; skips passes whose digit is
 * constant across the input. */
void or_sort_u64(uint64_t* a, uint64_t n) {
    if (n < 2) return;
    uint64_t* tmp = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint64_t* src = a;
    uint64_t* dst = tmp;
    static const int kBits = 16;
    uint64_t* cnt = (uint64_t*)malloc(((size_t)1 << kBits) * sizeof(uint64_t));
    for (int pass = 0; pass < 4; pass++) {
        int sh = pass * kBits;
        memset(cnt, 0, ((size_t)1 << kBits) * sizeof(uint64_t));
        for (uint64_t i = 0; i < n; i++) cnt[(src[i] >> sh) & 0xFFFF]++;
        if (cnt[(src[0] >> sh) & 0xFFFF] == n) continue;
        uint64_t s = 0;
        for (size_t d = 0; d < ((size_t)1 << kBits); d++) {
            uint64_t c = cnt[d];
            cnt[d] = s;
            s += c;
        }
        for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 0xFFFF]++] = src[i];
        uint64_t* t = src;
        src = dst;
        dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(uint64_t));
    free(cnt);
    free(tmp);
}

or_index* or_index_new(uint32_t ndir, uint32_t ebits) {
    or_index* x = (or_index*)calloc(1, sizeof(or_index));
    x->ndir = ndir;
    x->ebits = ebits;
    x->ushift = 64 - ebits;
    x->umask = (1u << (32 - ebits)) - 1u;
    x->upd = (or_upd*)calloc(ndir, sizeof(or_upd));
    return x;
}

void or_index_free(or_index* x) {
    if (!x) return;
    for (uint64_t i = 0; i < (uint64_t)x->dv_cap * x->ndir; i++) free(x->m[i].d);
    free(x->m);
    for (uint32_t e = 0; e < x->ndir; e++) {
        free(x->upd[e].keys);
        free(x->upd[e].off);
        free(x->upd[e].data);
    }
    free(x->upd);
    free(x);
}

void or_index_set_dv(or_index* x, uint32_t dv) {
    if (dv > x->dv_cap) {
        uint32_t nc = x->dv_cap ? x->dv_cap : 1024;
        while (nc < dv) nc *= 2;
        x->m = (or_row*)realloc(x->m, (size_t)nc * x->ndir * sizeof(or_row));
        memset(x->m + (size_t)x->dv_cap * x->ndir, 0, (size_t)(nc - x->dv_cap) * x->ndir * sizeof(or_row));
        x->dv_cap = nc;
    }
    if (dv > x->dv) x->dv = dv;
}

uint32_t or_index_dv(const or_index* x) { return x->dv; }

static int check_keys(or_index* x, const uint64_t* k, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) {
        uint32_t e = (uint32_t)(k[i] >> x->ushift), u = (uint32_t)(k[i] >> 32) & x->umask, v = (uint32_t)k[i];
        if (e >= x->ndir) return 1;
        if (u >= x->dv || v >= x->dv) {
            uint32_t m = u > v ? u : v;
            or_index_set_dv(x, m + 1);
        }
    }
    return 0;
}

static uint64_t* sorted_copy(const uint64_t* k, uint64_t n, int uniq, uint64_t* n_out) {
    uint64_t* s = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    if (n) memcpy(s, k, n * sizeof(uint64_t));
    or_sort_u64(s, n);
    uint64_t m = n;
    if (uniq && n) {
        m = 1;
        for (uint64_t i = 1; i < n; i++)
            if (s[i] != s[m - 1]) s[m++] = s[i];
    }
    *n_out = m;
    return s;
}

/* Apply a sorted delta group by group: insert = set union, remove = set difference. */
static void apply_sorted(or_index* x, const uint64_t* s, uint64_t n, int insert) {
    uint32_t* tmp = NULL;
    uint32_t tcap = 0;
    uint64_t i = 0;
    while (i < n) {
        uint64_t hi = s[i] >> 32;
        uint64_t j = i;
        while (j < n && (s[j] >> 32) == hi) j++;
        uint32_t e = (uint32_t)(s[i] >> x->ushift), u = (uint32_t)hi & x->umask;
        or_row* r = &x->m[(uint64_t)u * x->ndir + e];
        uint32_t need = r->len + (uint32_t)(j - i);
        if (need > tcap) {
            tcap = need * 2;
            tmp = (uint32_t*)realloc(tmp, (size_t)tcap * sizeof(uint32_t));
        }
        uint32_t a = 0, c = 0;
        uint64_t b = i;
        if (insert) {
            while (a < r->len || b < j) {
                uint32_t bv = b < j ? (uint32_t)s[b] : 0;
                if (b >= j || (a < r->len && r->d[a] < bv)) tmp[c++] = r->d[a++];
                else if (a >= r->len || bv < r->d[a]) {
                    tmp[c++] = bv;
                    b++;
                } else {
                    tmp[c++] = r->d[a++];
                    b++;
                }
            }
        } else {
            while (a < r->len) {
                while (b < j && (uint32_t)s[b] < r->d[a]) b++;
                if (b < j && (uint32_t)s[b] == r->d[a]) a++;
                else tmp[c++] = r->d[a++];
            }
        }
        x->m_entries += (uint64_t)c - r->len;
        x->last_touched_bytes += 4ull * (r->len + (j - i) + c);
        row_reserve(r, c);
        if (c) memcpy(r->d, tmp, (size_t)c * sizeof(uint32_t));
        r->len = c;
        i = j;
    }
    free(tmp);
}

int or_index_init(or_index* x, const uint64_t* keys, uint64_t n) {
    if (check_keys(x, keys, n)) return 1;
    for (uint64_t i = 0; i < (uint64_t)x->dv_cap * x->ndir; i++) x->m[i].len = 0;
    x->m_entries = 0;
    return or_index_insert(x, keys, n);
}

int or_index_insert(or_index* x, const uint64_t* keys, uint64_t n) {
    if (check_keys(x, keys, n)) return 1;
    uint64_t m;
    uint64_t* s = sorted_copy(keys, n, 1, &m);
    x->last_touched_bytes = 0;
    apply_sorted(x, s, m, 1);
    free(s);
    return 0;
}

int or_index_remove(or_index* x, const uint64_t* keys, uint64_t n) {
    if (check_keys(x, keys, n)) return 1;
    uint64_t m;
    uint64_t* s = sorted_copy(keys, n, 0, &m);
    x->last_touched_bytes = 0;
    apply_sorted(x, s, m, 0);
    free(s);
    return 0;
}

int or_index_rebuild_update(or_index* x, const uint64_t* keys, uint64_t n) {
    if (check_keys(x, keys, n)) return 1;
    uint64_t m;
    uint64_t* s = sorted_copy(keys, n, 0, &m);
    for (uint32_t e = 0; e < x->ndir; e++) {
        or_upd* U = &x->upd[e];
        U->nkeys = 0;
        U->ndata = 0;
    }
    x->u_entries = m;
    uint64_t i = 0;
    while (i < m) {
        uint32_t e = (uint32_t)(s[i] >> x->ushift);
        uint64_t j = i;
        while (j < m && (uint32_t)(s[j] >> x->ushift) == e) j++;
        or_upd* U = &x->upd[e];
        uint64_t cnt = j - i, nk = 0;
        for (uint64_t k = i; k < j; k++)
            if (k == i || (s[k] >> 32) != (s[k - 1] >> 32)) nk++;
        U->keys = (uint32_t*)realloc(U->keys, (nk ? nk : 1) * sizeof(uint32_t));
        U->off = (uint64_t*)realloc(U->off, (nk + 1) * sizeof(uint64_t));
        U->data = (uint32_t*)realloc(U->data, (cnt ? cnt : 1) * sizeof(uint32_t));
        nk = 0;
        for (uint64_t k = i; k < j; k++) {
            if (k == i || (s[k] >> 32) != (s[k - 1] >> 32)) {
                U->keys[nk] = (uint32_t)(s[k] >> 32) & x->umask;
                U->off[nk] = k - i;
                nk++;
            }
            U->data[k - i] = (uint32_t)s[k];
        }
        U->off[nk] = cnt;
        U->nkeys = nk;
        U->ndata = cnt;
        i = j;
    }
    free(s);
    return 0;
}

uint64_t or_index_entries(const or_index* x, int which) { return which ? x->u_entries : x->m_entries; }

uint64_t or_index_export(const or_index* x, int which, uint64_t* out, uint64_t cap) {
    uint64_t p = 0;
    for (uint32_t e = 0; e < x->ndir; e++) {
        if (which == 0) {
            for (uint32_t u = 0; u < x->dv; u++) {
                const or_row* r = &x->m[(uint64_t)u * x->ndir + e];
                for (uint32_t k = 0; k < r->len; k++, p++)
                    if (p < cap) out[p] = ((uint64_t)e << x->ushift) | ((uint64_t)u << 32) | r->d[k];
            }
        } else {
            const or_upd* U = &x->upd[e];
            for (uint64_t k = 0; k < U->nkeys; k++)
                for (uint64_t t = U->off[k]; t < U->off[k + 1]; t++, p++)
                    if (p < cap) out[p] = ((uint64_t)e << x->ushift) | ((uint64_t)U->keys[k] << 32) | U->data[t];
        }
    }
    return p;
}

const uint32_t* or_upd_list(const or_index* x, uint32_t e, uint32_t u, uint32_t* len) {
    const or_upd* U = &x->upd[e];
    uint64_t lo = 0, hi = U->nkeys;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (U->keys[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    if (lo < U->nkeys && U->keys[lo] == u) {
        *len = (uint32_t)(U->off[lo + 1] - U->off[lo]);
        return U->data + U->off[lo];
    }
    *len = 0;
    return NULL;
}
