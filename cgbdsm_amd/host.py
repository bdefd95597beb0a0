"""Python view of the host pipeline (Stage-1 CaLiG producer, planner, key
encoder) through include/cgbdsm_host.h.  Mirrors the batch loop of
main_calig.cpp:216-275: each `next_batch()` runs updateIndex ->
ConstructUpdate -> mergeDeltas -> PrepareUpdateAll -> PrepareUpdateIndex and
returns the keys handed to Stage 2."""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N

QUERY_KINDS = {"tree": 0, "dense": 1, "sparse": 2}


def _check(rc, lib):
    if rc != 0:
        raise RuntimeError(lib.cgh_last_error().decode())


def synth_spec(scale, avg_degree, num_labels, zipf_s=0.0, initial_frac=0.9, stream_edges=0, seed=1,
               query_size=5, query_kind="tree", query_seed=1, a=0.57, b=0.19, c=0.19):
    s = N.SynthSpec()
    s.scale, s.avg_degree, s.num_labels, s.zipf_s = scale, avg_degree, num_labels, zipf_s
    s.a, s.b, s.c = a, b, c
    s.initial_frac, s.stream_edges, s.seed = initial_frac, stream_edges, seed
    s.query_size, s.query_kind, s.query_seed = query_size, QUERY_KINDS[query_kind], query_seed
    return s


def generate_files(spec, query_path, data_path, update_path):
    L = N.host_lib()
    _check(L.cgh_generate_files(C.byref(spec), str(query_path).encode(), str(data_path).encode(),
                                str(update_path).encode()), L)


def generate_binary(spec, query_path, prefix):
    """Query text file + the graph as a binary CSR prefix (update stream
    split off by the reader)."""
    L = N.host_lib()
    _check(L.cgh_generate_binary(C.byref(spec), str(query_path).encode(), str(prefix).encode()), L)


@dataclass
class KeyBatch:
    index: int
    ins: np.ndarray
    dele: np.ndarray
    upd: np.ndarray
    edge_ok: np.ndarray
    dv: int
    pairs: int
    s1_ms: dict


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class HostSession:
    """One query over one data graph and its update stream."""

    def __init__(self, *, files=None, spec=None, batch_edges=None, batch_ints=None, threads=0, ebits=5):
        """batch_edges: stream edges per batch; batch_ints: the same in ints
        (two per edge) as the reference CLI counts them, odd values included
        (main_calig.cpp:65-69).  Neither: one batch."""
        L = self._L = N.host_lib()
        h = C.c_void_p()
        if batch_ints is None:
            batch_ints = 0 if batch_edges is None else 2 * int(batch_edges)
        if files is not None:
            q, d, u = files
            _check(L.cgh_session_open(str(q).encode(), str(d).encode(), str(u).encode(), batch_ints,
                                      threads, ebits, C.byref(h)), L)
        else:
            _check(L.cgh_session_open_synth(C.byref(spec), batch_ints, threads, ebits, C.byref(h)), L)
        self._h = h
        self.ebits = ebits
        self.query_tables = N.Plan()
        _check(L.cgh_query_tables(h, C.byref(self.query_tables)), L)
        self.qv, self.qe = self.query_tables.qv, self.query_tables.qe

    def close(self):
        if self._h:
            self._L.cgh_session_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_batches(self):
        return self._L.cgh_num_batches(self._h)

    @property
    def threads(self):
        return self._L.cgh_num_threads(self._h)

    def initial_keys(self):
        L = self._L
        ptr = C.POINTER(C.c_uint64)()
        n, dv = C.c_uint64(), C.c_uint32()
        card = (C.c_uint32 * N.MAX_DIR)()
        avg = (C.c_float * N.MAX_DIR)()
        _check(L.cgh_initial_keys(self._h, C.byref(ptr), C.byref(n), C.byref(dv), card, avg), L)
        nd = 2 * self.qe
        return (_arr(ptr, n.value, np.uint64), dv.value, np.array(card[:nd], dtype=np.uint32),
                np.array(avg[:nd], dtype=np.float32))

    def make_plan(self, card, avg):
        L = self._L
        p = N.Plan()
        c = (C.c_uint32 * N.MAX_DIR)(*[int(x) for x in card])
        a = (C.c_float * N.MAX_DIR)(*[float(x) for x in avg])
        _check(L.cgh_make_plan(self._h, c, a, C.byref(p)), L)
        return p

    def stream_edges_left(self):
        return int(self._L.cgh_stream_left(self._h)) // 2

    def next_batch(self, n_edges=None):
        """Stage 1 of the next batch: the session's fixed batch size, or the
        next `n_edges` stream edges (batches of any size)."""
        L = self._L
        b = N.Batch()
        if n_edges is None:
            _check(L.cgh_next_batch(self._h, C.byref(b)), L)
        else:
            _check(L.cgh_next_batch_edges(self._h, int(n_edges), C.byref(b)), L)
        return KeyBatch(b.batch, _arr(b.ins, b.n_ins, np.uint64), _arr(b.dele, b.n_del, np.uint64),
                        _arr(b.upd, b.n_upd, np.uint64), _arr(b.edge_ok, self.qe, np.uint8), b.dv, b.pairs,
                        {"updateIndex": b.ms_update_index, "ConstructUpdate": b.ms_construct_update,
                         "PrepareUpdateAll": b.ms_prepare_all, "PrepareUpdateIndex": b.ms_prepare_index})

    def update_stream(self):
        """The loaded update stream as (n, 2) original-id pairs, in order."""
        L = self._L
        n = C.c_uint64()
        _check(L.cgh_export_updates(self._h, None, 0, C.byref(n)), L)
        out = np.zeros(max(n.value, 1), dtype=np.uint32)
        _check(L.cgh_export_updates(self._h, out.ctypes.data, n.value, C.byref(n)), L)
        return out[: n.value].reshape(-1, 2)

    def old_ids(self):
        ptr = C.POINTER(C.c_uint32)()
        n = C.c_uint32()
        _check(self._L.cgh_old_ids(self._h, C.byref(ptr), C.byref(n)), self._L)
        return _arr(ptr, n.value, np.uint32)

    def graph(self):
        """Current label-filtered data graph as CSR (offs, adj, labels)."""
        L = self._L
        n, m = C.c_uint32(), C.c_uint64()
        _check(L.cgh_export_graph(self._h, None, None, 0, None, C.byref(n), C.byref(m)), L)
        offs = np.zeros(n.value + 1, dtype=np.uint64)
        adj = np.zeros(max(m.value, 1), dtype=np.uint32)
        lab = np.zeros(n.value, dtype=np.int32)
        _check(L.cgh_export_graph(self._h, offs.ctypes.data, adj.ctypes.data, m.value, lab.ctypes.data,
                                  C.byref(n), C.byref(m)), L)
        return offs, adj, lab

    def lit(self, n):
        li = np.zeros(n, dtype=np.uint32)
        _check(self._L.cgh_export_li(self._h, li.ctypes.data, n), self._L)
        return li

    def cand(self, v, ui, uj, cap=1 << 20):
        out = np.zeros(cap, dtype=np.uint32)
        n = C.c_int64()
        _check(self._L.cgh_export_cand(self._h, v, ui, uj, out.ctypes.data, cap, C.byref(n)), self._L)
        return None if n.value < 0 else out[: n.value].copy()

    def query_info(self):
        lab = np.zeros(N.MAX_QV, dtype=np.int32)
        ed = np.zeros(2 * N.MAX_QE, dtype=np.uint32)
        qv, qe = C.c_uint32(), C.c_uint32()
        _check(self._L.cgh_query_info(self._h, lab.ctypes.data, ed.ctypes.data, C.byref(qv), C.byref(qe)),
               self._L)
        return lab[: qv.value].copy(), ed[: 2 * qe.value].copy()


class Workload:
    """A synthetic graph and its update stream without any host CaLiG (for
    the device Stage 1): the same graph a HostSession of the spec builds."""

    def __init__(self, spec):
        L = self._L = N.host_lib()
        self._h = C.c_void_p()
        _check(L.cgh_workload_open(C.byref(spec), C.byref(self._h)), L)
        n, m = C.c_uint32(), C.c_uint64()
        lab = C.POINTER(C.c_int32)()
        ed = C.POINTER(C.c_uint32)()
        _check(L.cgh_workload_graph(self._h, C.byref(n), C.byref(lab), C.byref(ed), C.byref(m)), L)
        self.n = n.value
        self.labels = _arr(lab, n.value, np.int32)
        self.edges = _arr(ed, 2 * m.value, np.uint32).reshape(-1, 2)
        ints = C.POINTER(C.c_int32)()
        ni = C.c_uint64()
        _check(L.cgh_workload_stream(self._h, C.byref(ints), C.byref(ni)), L)
        self.stream = _arr(ints, ni.value, np.int32)

    def query(self, size, kind, seed):
        """(labels, edges) of a seeded random-walk query of `size` vertices."""
        lab = np.zeros(N.MAX_QV, dtype=np.int32)
        ed = np.zeros(2 * N.MAX_QE, dtype=np.uint32)
        qv, qe = C.c_uint32(), C.c_uint32()
        _check(self._L.cgh_workload_query(self._h, size, QUERY_KINDS[kind], seed, lab.ctypes.data, ed.ctypes.data,
                                          C.byref(qv), C.byref(qe)), self._L)
        return lab[: qv.value].copy(), ed[: 2 * qe.value].reshape(-1, 2).copy()

    def close(self):
        if self._h:
            self._L.cgh_workload_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Planner:
    """The planner of one query given by labels and edges (no CaLiG)."""

    def __init__(self, labels, edges, ebits=5):
        L = self._L = N.host_lib()
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        edges = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1))
        self._h = C.c_void_p()
        _check(L.cgh_planner_open(labels.ctypes.data, labels.size, edges.ctypes.data, edges.size // 2, ebits,
                                  C.byref(self._h)), L)
        self.labels = labels
        self.ebits = ebits
        self.query_tables = N.Plan()
        _check(L.cgh_planner_tables(self._h, C.byref(self.query_tables)), L)
        self.qv, self.qe = self.query_tables.qv, self.query_tables.qe

    def make_plan(self, card, avg):
        p = N.Plan()
        c = (C.c_uint32 * N.MAX_DIR)(*[int(x) for x in card])
        a = (C.c_float * N.MAX_DIR)(*[float(x) for x in avg])
        _check(self._L.cgh_planner_make_plan(self._h, c, a, C.byref(p)), self._L)
        return p

    def close(self):
        if self._h:
            self._L.cgh_planner_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
