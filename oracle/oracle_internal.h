/* Internal layout of the oracle index -- TEST INFRASTRUCTURE ONLY. */
#ifndef CGBDSM_ORACLE_INTERNAL_H
#define CGBDSM_ORACLE_INTERNAL_H
#include <stdint.h>

typedef struct {
    uint32_t* d;
    uint32_t len, cap;
} or_row;

typedef struct {
    uint32_t* keys;   /* sorted owners with a list   */
    uint64_t* off;    /* nkeys + 1                    */
    uint32_t* data;
    uint64_t nkeys, ndata;
} or_upd;

struct or_index {
    uint32_t ndir, ebits, ushift, umask;
    uint32_t dv, dv_cap;
    or_row* m;        /* [owner * ndir + dir] */
    or_upd* upd;      /* [dir]                */
    uint64_t m_entries, u_entries, last_touched_bytes;
};

void or_sort_u64(uint64_t* a, uint64_t n);
const uint32_t* or_upd_list(const struct or_index* x, uint32_t e, uint32_t u, uint32_t* len);

static inline const uint32_t* or_main_list(const struct or_index* x, uint32_t e, uint32_t u, uint32_t* len) {
    if (u >= x->dv) {
        *len = 0;
        return 0;
    }
    const or_row* r = &x->m[(uint64_t)u * x->ndir + e];
    *len = r->len;
    return r->d;
}

#endif
