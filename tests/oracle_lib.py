"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test
infrastructure only; the product path never imports this module."""
import ctypes as C
import os

import numpy as np

from cgbdsm_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise OSError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
        L.or_index_new.argtypes = [U32, U32]
        L.or_index_new.restype = P
        L.or_index_free.argtypes = [P]
        L.or_index_free.restype = None
        L.or_index_set_dv.argtypes = [P, U32]
        L.or_index_set_dv.restype = None
        L.or_index_dv.argtypes = [P]
        L.or_index_dv.restype = U32
        for f in ("or_index_init", "or_index_insert", "or_index_remove", "or_index_rebuild_update"):
            getattr(L, f).argtypes = [P, C.c_void_p, U64]
            getattr(L, f).restype = I
        L.or_index_export.argtypes = [P, I, C.c_void_p, U64]
        L.or_index_export.restype = U64
        L.or_index_entries.argtypes = [P, I]
        L.or_index_entries.restype = U64
        L.or_masks.argtypes = [P, C.POINTER(N.Plan), C.c_void_p, C.c_void_p]
        L.or_mask_words.argtypes = [P]
        L.or_mask_words.restype = U64
        L.or_match.argtypes = [P, C.POINTER(N.Plan), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_void_p, U64, I]
        L.or_match.restype = U64
        L.or_brute_count.argtypes = [U32, C.c_void_p, C.c_void_p, C.c_void_p, U32, C.c_void_p, U32, C.c_void_p]
        L.or_brute_count.restype = U64
        L.or_anchored_count.argtypes = [U32, C.c_void_p, C.c_void_p, C.c_void_p, U32, C.c_void_p, U32,
                                        C.c_void_p, C.c_void_p, U64, C.c_void_p, I]
        L.or_anchored_count.restype = U64
        L.or_li_fixpoint.argtypes = [U32, C.c_void_p, C.c_void_p, C.c_void_p, U32, C.c_void_p, U32, C.c_void_p,
                                     C.c_void_p]
        L.or_brute_tuples.argtypes = [U32, C.c_void_p, C.c_void_p, C.c_void_p, U32, C.c_void_p, U32, C.c_void_p,
                                      C.c_void_p, U64, C.c_void_p, C.c_void_p, U64]
        L.or_brute_tuples.restype = U64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


class Index:
    """Set-semantics mirror of the main index M and update index U."""

    def __init__(self, ndir, ebits):
        self.L = lib()
        self.h = self.L.or_index_new(ndir, ebits)
        self.ndir = ndir

    def __del__(self):
        try:
            self.L.or_index_free(self.h)
        except Exception:
            pass

    def set_dv(self, dv):
        self.L.or_index_set_dv(self.h, dv)

    @property
    def dv(self):
        return self.L.or_index_dv(self.h)

    def _op(self, f, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        assert getattr(self.L, f)(self.h, _p(keys), keys.size) == 0

    def init(self, keys):
        self._op("or_index_init", keys)

    def insert(self, keys):
        self._op("or_index_insert", keys)

    def remove(self, keys):
        self._op("or_index_remove", keys)

    def rebuild_update(self, keys):
        self._op("or_index_rebuild_update", keys)

    def export(self, which):
        n = self.L.or_index_export(self.h, which, None, 0)
        out = np.zeros(max(n, 1), dtype=np.uint64)
        self.L.or_index_export(self.h, which, out.ctypes.data, n)
        return out[:n]

    def masks(self, plan):
        W = self.L.or_mask_words(self.h)
        out = np.zeros(max(plan.qe * 12 * W, 1), dtype=np.uint64)
        ones = np.zeros(plan.qe * 12, dtype=np.uint8)
        self.L.or_masks(self.h, C.byref(plan), out.ctypes.data, ones.ctypes.data)
        return out, ones, W

    def match(self, plan, edge_ok, masks=None, tuples_cap=0, threads=1):
        per = np.zeros(plan.qe, dtype=np.uint64)
        eok = np.ascontiguousarray(edge_ok, dtype=np.uint8)
        m, ones = (None, None) if masks is None else (masks[0], masks[1])
        tup = np.zeros(max(tuples_cap * plan.qv, 1), dtype=np.uint32) if tuples_cap else None
        tei = np.zeros(max(tuples_cap, 1), dtype=np.uint8) if tuples_cap else None
        total = self.L.or_match(self.h, C.byref(plan), _p(m), _p(ones), eok.ctypes.data, per.ctypes.data,
                                _p(tup), _p(tei), tuples_cap, threads)
        if tuples_cap:
            n = min(total, tuples_cap)
            return total, per, tup[: n * plan.qv].reshape(n, plan.qv), tei[:n]
        return total, per


def brute_count(offs, adj, labels, qlabels, qedges):
    L = lib()
    qlabels = np.ascontiguousarray(qlabels, dtype=np.int32)
    qedges = np.ascontiguousarray(qedges, dtype=np.uint32)
    return L.or_brute_count(len(labels), offs.ctypes.data, adj.ctypes.data, labels.ctypes.data, qlabels.size,
                            qlabels.ctypes.data, qedges.size // 2, qedges.ctypes.data)


def anchored_count(offs, adj, labels, qlabels, qedges, batch_pairs, threads=4):
    """batch_pairs: (k,2) array of the batch's edges."""
    L = lib()
    qlabels = np.ascontiguousarray(qlabels, dtype=np.int32)
    qedges = np.ascontiguousarray(qedges, dtype=np.uint32)
    bp = np.asarray(batch_pairs, dtype=np.uint64).reshape(-1, 2)
    keys = np.sort((np.minimum(bp[:, 0], bp[:, 1]) << np.uint64(32)) | np.maximum(bp[:, 0], bp[:, 1]))
    keys = np.unique(keys)
    per = np.zeros(qedges.size // 2, dtype=np.uint64)
    tot = L.or_anchored_count(len(labels), offs.ctypes.data, adj.ctypes.data, labels.ctypes.data, qlabels.size,
                              qlabels.ctypes.data, qedges.size // 2, qedges.ctypes.data, _p(keys), keys.size,
                              per.ctypes.data, threads)
    return tot, per


def _batch_keys(batch_pairs):
    bp = np.asarray(batch_pairs, dtype=np.uint64).reshape(-1, 2)
    return np.unique((np.minimum(bp[:, 0], bp[:, 1]) << np.uint64(32)) | np.maximum(bp[:, 0], bp[:, 1]))


def brute_tuples(offs, adj, labels, qlabels, qedges, batch_pairs, cap=1 << 20):
    """New embeddings of one batch (graph = G_b) as an (n, qv) array of
    original ids in query-vertex order, and the tag of each."""
    L = lib()
    qlabels = np.ascontiguousarray(qlabels, dtype=np.int32)
    qedges = np.ascontiguousarray(qedges, dtype=np.uint32)
    keys = _batch_keys(batch_pairs)
    qv = qlabels.size
    out = np.zeros(max(cap * qv, 1), dtype=np.uint32)
    tags = np.zeros(max(cap, 1), dtype=np.uint8)
    n = L.or_brute_tuples(len(labels), offs.ctypes.data, adj.ctypes.data, labels.ctypes.data, qv, qlabels.ctypes.data,
                          qedges.size // 2, qedges.ctypes.data, _p(keys), keys.size, out.ctypes.data,
                          tags.ctypes.data, cap)
    if n > cap:
        raise ValueError(f"{n} new embeddings exceed the tuple capacity {cap}")
    return out[: n * qv].reshape(n, qv), tags[:n]


def li_fixpoint(offs, adj, labels, qlabels, qedges):
    L = lib()
    qlabels = np.ascontiguousarray(qlabels, dtype=np.int32)
    qedges = np.ascontiguousarray(qedges, dtype=np.uint32)
    li = np.zeros(len(labels), dtype=np.uint32)
    L.or_li_fixpoint(len(labels), offs.ctypes.data, adj.ctypes.data, labels.ctypes.data, qlabels.size,
                     qlabels.ctypes.data, qedges.size // 2, qedges.ctypes.data, li.ctypes.data)
    return li
