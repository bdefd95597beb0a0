"""Python handle on the MI355X Stage-2 engine (include/cgbdsm.h).

`Engine` owns a device context; `DeviceQuery` mirrors the reference's
GPUGraphLoader + CaLiGHelperGamma device surface
(graph/graph_gpu.h:152-188, index/calig_helper_gamma.h:91-110):

    q = eng.query(query_tables)                 # ctor + LoadQuery
    card, avg = q.index_init(keys, dv)          # ConvertGlobalIndex
    q.set_plan(plan)                            # LoadPlan
    q.stage(batch)                              # staging of one batch
    q.transfer_update_all()                     # TransferUpdateAll
    q.transfer_update_index()                   # TransferUpdateIndex
    q.build_local_index()                       # BuildLocalIndexBitAll
    count, per_edge = q.matching()              # MatchingBitAll
    q.batch_update_global_index()               # BatchUpdateGlobalIndex

There is no host fallback: a missing libcgbdsm.so raises at import of the
handle, and every non-zero status raises `CgbdsmError`.
"""
import ctypes as C

import numpy as np

from . import _native as N

STATUS = {0: "OK", 1: "INVALID", 2: "OOM", 3: "DEVICE", 4: "LIMIT", 5: "OVERFLOW"}


class CgbdsmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise CgbdsmError(rc, N.dev_lib().cgbdsm_last_error().decode())


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


class Engine:
    def __init__(self, device=0, budget_bytes=0):
        self.L = N.dev_lib()
        self.h = C.c_void_p()
        _check(self.L.cgbdsm_ctx_create(device, budget_bytes, C.byref(self.h)))

    def query(self, tables):
        return DeviceQuery(self, tables)

    def memory(self):
        """(device bytes held by this engine's queries, budget or 0)."""
        used, budget = C.c_uint64(), C.c_uint64()
        _check(self.L.cgbdsm_ctx_memory(self.h, C.byref(used), C.byref(budget)))
        return int(used.value), int(budget.value)

    def graph(self, labels, edges):
        """Device data graph shared by this engine's queries (device Stage 1)."""
        return DeviceGraph(self, labels, edges)

    def matching_batch(self, queries):
        """Matching of several queries of this engine (each after
        build_local_index) with one final-DFS launch; returns a list of
        (count, per_edge) in the order given."""
        n = len(queries)
        hs = (C.c_void_p * n)(*[q.h.value for q in queries])
        counts = (C.c_ulonglong * n)()
        pers = [np.zeros(N.MAX_QE, dtype=np.uint64) for _ in queries]
        pp = (C.c_void_p * n)(*[p.ctypes.data for p in pers])
        _check(self.L.cgbdsm_matching_batch(hs, n, counts, pp))
        return [(int(counts[i]), pers[i][: q.qe].copy()) for i, q in enumerate(queries)]

    def close(self):
        if self.h:
            _check(self.L.cgbdsm_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """CaLiG's data graph in HBM (cgbdsm_graph_*): labels, adjacency, the
    update stream and the latest batch's new edges."""

    def __init__(self, engine, labels, edges):
        self.eng, self.L = engine, engine.L
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        edges = np.ascontiguousarray(np.asarray(edges, dtype=np.uint32).reshape(-1, 2))
        self.n = int(labels.size)
        self.h = C.c_void_p()
        _check(self.L.cgbdsm_graph_create(engine.h, self.n, labels.ctypes.data, _ptr(edges), edges.shape[0],
                                          C.byref(self.h)))

    def set_stream(self, ints):
        ints = np.ascontiguousarray(np.asarray(ints).reshape(-1), dtype=np.int32)
        _check(self.L.cgbdsm_graph_set_stream(self.h, _ptr(ints), ints.size))

    def apply_range(self, start, count):
        _check(self.L.cgbdsm_graph_apply_range(self.h, int(start), int(count)))

    def info(self):
        b, m, nn = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(self.L.cgbdsm_graph_info(self.h, C.byref(b), C.byref(m), C.byref(nn)))
        return {"batches": b.value, "directed_edges": m.value, "last_new": nn.value}

    def close(self):
        if self.h:
            _check(self.L.cgbdsm_graph_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceQuery:
    def __init__(self, engine, tables):
        self.eng, self.L = engine, engine.L
        self.h = C.c_void_p()
        self.tables = tables
        self.qv, self.qe, self.ndir = tables.qv, tables.qe, tables.ndir
        _check(self.L.cgbdsm_query_create(engine.h, C.byref(tables), C.byref(self.h)))

    def close(self):
        if self.h:
            _check(self.L.cgbdsm_query_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_plan(self, plan):
        _check(self.L.cgbdsm_query_set_plan(self.h, C.byref(plan)))
        self.tables = plan

    def index_init(self, keys, dv):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        card = np.zeros(N.MAX_DIR, dtype=np.uint32)
        avg = np.zeros(N.MAX_DIR, dtype=np.float32)
        _check(self.L.cgbdsm_index_init(self.h, _ptr(keys), keys.size, dv, card.ctypes.data, avg.ctypes.data))
        return card[: self.ndir], avg[: self.ndir]

    def calig_attach(self, graph, qlabels):
        """Device Stage 1: CaLiG index of this query over `graph`, then the
        initial main index; returns (cardinalities, avg_degrees)."""
        ql = np.zeros(N.MAX_QV, dtype=np.int32)
        ql[: len(qlabels)] = qlabels
        card = np.zeros(N.MAX_DIR, dtype=np.uint32)
        avg = np.zeros(N.MAX_DIR, dtype=np.float32)
        _check(self.L.cgbdsm_calig_attach(self.h, graph.h, ql.ctypes.data, card.ctypes.data, avg.ctypes.data))
        self.graph = graph
        return card[: self.ndir], avg[: self.ndir]

    def calig_update(self):
        """Device Stage 1 of the graph's latest batch; stages it."""
        _check(self.L.cgbdsm_calig_update(self.h))

    def calig_match_deleted(self):
        """Embeddings the staged batch's deletions destroy: (count, per_edge),
        (0, zeros) without deletions.  Call right after calig_update."""
        n = C.c_ulonglong(0)
        per = np.zeros(N.MAX_QE, dtype=np.uint64)
        _check(self.L.cgbdsm_calig_match_deleted(self.h, C.byref(n), per.ctypes.data))
        return int(n.value), per[: self.qe].copy()

    def calig_lit(self):
        li = np.zeros(self.graph.n, dtype=np.uint32)
        _check(self.L.cgbdsm_calig_export_lit(self.h, li.ctypes.data, self.graph.n))
        return li

    def calig_old_ids(self):
        n = C.c_uint32()
        _check(self.L.cgbdsm_calig_old_ids(self.h, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint32)
        _check(self.L.cgbdsm_calig_old_ids(self.h, out.ctypes.data, n.value, C.byref(n)))
        return out[: n.value]

    def calig_stats(self):
        ms = C.c_double()
        o = np.zeros(6, dtype=np.uint64)
        _check(self.L.cgbdsm_calig_stats(self.h, C.byref(ms), o.ctypes.data))
        return {"ms": ms.value, "rounds_act": int(o[0]), "rounds_del": int(o[1]), "activated": int(o[2]),
                "killed": int(o[3]), "new_pairs": int(o[4]), "dv": int(o[5])}

    def staged_keys(self, which):
        """Keys of the staged batch: 'ins', 'del' or 'upd' (unsorted)."""
        w = {"ins": 0, "del": 1, "upd": 2}[which]
        n = C.c_uint64()
        _check(self.L.cgbdsm_staged_keys(self.h, w, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(self.L.cgbdsm_staged_keys(self.h, w, out.ctypes.data, n.value, C.byref(n)))
        return out[: n.value]

    def stage(self, kb):
        ins = np.ascontiguousarray(kb.ins, dtype=np.uint64)
        dele = np.ascontiguousarray(kb.dele, dtype=np.uint64)
        upd = np.ascontiguousarray(kb.upd, dtype=np.uint64)
        eok = np.ascontiguousarray(kb.edge_ok, dtype=np.uint8)
        _check(self.L.cgbdsm_batch_stage(self.h, _ptr(ins), ins.size, _ptr(dele), dele.size, _ptr(upd), upd.size,
                                         eok.ctypes.data, kb.dv))

    def enqueue(self, kb):
        """Copy one batch's keys into the device-resident FIFO."""
        ins = np.ascontiguousarray(kb.ins, dtype=np.uint64)
        dele = np.ascontiguousarray(kb.dele, dtype=np.uint64)
        upd = np.ascontiguousarray(kb.upd, dtype=np.uint64)
        eok = np.ascontiguousarray(kb.edge_ok, dtype=np.uint8)
        _check(self.L.cgbdsm_batch_enqueue(self.h, _ptr(ins), ins.size, _ptr(dele), dele.size, _ptr(upd),
                                           upd.size, eok.ctypes.data, kb.dv))

    def stage_next(self):
        st = C.c_int()
        _check(self.L.cgbdsm_batch_stage_next(self.h, C.byref(st)))
        return bool(st.value)

    def transfer_update_all(self):
        _check(self.L.cgbdsm_transfer_update_all(self.h))

    def transfer_update_index(self):
        ne = C.c_int()
        _check(self.L.cgbdsm_transfer_update_index(self.h, C.byref(ne)))
        return bool(ne.value)

    def build_local_index(self):
        _check(self.L.cgbdsm_build_local_index(self.h))

    def build_frontier(self):
        """Seeds and BFS levels of the next matching (optional; see cgbdsm_build_frontier)."""
        _check(self.L.cgbdsm_build_frontier(self.h))

    def matching(self):
        n = C.c_ulonglong(0)
        per = np.zeros(N.MAX_QE, dtype=np.uint64)
        _check(self.L.cgbdsm_matching(self.h, C.byref(n), per.ctypes.data))
        return int(n.value), per[: self.qe].copy()

    def batch_update_global_index(self):
        _check(self.L.cgbdsm_batch_update_global_index(self.h))

    def run_batch(self):
        """Whole Stage 2 of a staged batch; returns (count, per_edge, skipped)."""
        n = C.c_ulonglong(0)
        per = np.zeros(N.MAX_QE, dtype=np.uint64)
        sk = C.c_int()
        _check(self.L.cgbdsm_run_batch(self.h, C.byref(n), per.ctypes.data, C.byref(sk)))
        return int(n.value), per[: self.qe].copy(), bool(sk.value)

    def set_match_params(self, bfs_row_limit=65536, frontier_cap=0, donate_min=4):
        _check(self.L.cgbdsm_set_match_params(self.h, bfs_row_limit, frontier_cap, donate_min))
        self.match_params = (bfs_row_limit, frontier_cap, donate_min)

    def set_layout(self, layout):
        """'gamma' (per-list merges) or 'gpma' (packed-memory array); before index_init."""
        _check(self.L.cgbdsm_query_set_layout(self.h, {"gamma": 0, "gpma": 1}[layout]))

    def set_dfs_options(self, shortest_list=True, product_suffix=True):
        """Count-mode refinements of the final DFS (both on by default)."""
        _check(self.L.cgbdsm_set_dfs_options(self.h, int(bool(shortest_list)), int(bool(product_suffix))))

    def set_dfs_tuning(self, rows_per_grab=4, blocks_per_cu=0):
        """Launch shape of the final DFS (counts are identical for every setting)."""
        _check(self.L.cgbdsm_set_dfs_tuning(self.h, int(rows_per_grab), int(blocks_per_cu)))
        self.dfs_tuning = (int(rows_per_grab), int(blocks_per_cu))

    def set_dfs_search(self, wave=True, min_len=64):
        """Wave-cooperative list search of the final DFS (on by default,
        min_len 64): lists of at most 64 entries held one per lane, lists of
        min_len entries or more (a power of two >= 64) searched by pivots."""
        _check(self.L.cgbdsm_set_dfs_search(self.h, int(min_len) if wave else 0))

    def set_dfs_steal(self, range_min=1024, cache_cap=0):
        """Stealing chunk and stable-cache size of the final DFS (counts are
        identical for every setting)."""
        _check(self.L.cgbdsm_set_dfs_steal(self.h, int(range_min), int(cache_cap)))
        self.dfs_steal = (int(range_min), int(cache_cap))

    def export(self, which):
        n = C.c_uint64()
        _check(self.L.cgbdsm_index_export(self.h, which, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(self.L.cgbdsm_index_export(self.h, which, out.ctypes.data, n.value, C.byref(n)))
        return out[: n.value]

    def mask(self, ei, qv, dv):
        words = (dv + 63) // 64
        out = np.zeros(max(words, 1), dtype=np.uint64)
        ones = C.c_int()
        _check(self.L.cgbdsm_mask_export(self.h, ei, qv, out.ctypes.data, words, C.byref(ones)))
        return bool(ones.value), out[:words]

    def tuples(self, cap):
        out = np.zeros(max(cap * self.qv, 1), dtype=np.uint32)
        tei = np.zeros(max(cap, 1), dtype=np.uint8)
        n = C.c_uint64()
        _check(self.L.cgbdsm_match_tuples(self.h, out.ctypes.data, cap, C.byref(n), tei.ctypes.data))
        k = min(n.value, cap)
        return n.value, out[: k * self.qv].reshape(k, self.qv), tei[:k]

    def stats(self):
        s = N.Stats()
        _check(self.L.cgbdsm_stats_get(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in N.Stats._fields_}

    def set_deadline(self, ms):
        _check(self.L.cgbdsm_set_deadline_ms(self.h, float(ms)))

    def set_cpu_split(self, threads=-1, ratio=0.0, alpha=0.3, min_rows=0):
        """CPU/GPU split of the final DFS (cgbdsm_set_cpu_split): the host
        counts a share of the frontier rows while the GPU counts the rest.
        threads 0 turns it off, -1 uses hardware threads / NUMA nodes."""
        _check(self.L.cgbdsm_set_cpu_split(self.h, int(threads), float(ratio), float(alpha), int(min_rows)))

    def split_stats(self):
        s = N.SplitStats()
        _check(self.L.cgbdsm_cpu_split_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in N.SplitStats._fields_}

    def mirror_export(self, which):
        """The host mirror's M (0) or U (1) in the boundary key format."""
        n = C.c_uint64()
        _check(self.L.cgbdsm_cpu_mirror_export(self.h, which, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(self.L.cgbdsm_cpu_mirror_export(self.h, which, out.ctypes.data, n.value, C.byref(n)))
        return out[: n.value]

    def profile(self, enable=True):
        _check(self.L.cgbdsm_profile(self.h, 1 if enable else 0))

    def kernel_times(self, reset=False):
        arr = (N.KTime * len(N.KERNEL_CLASSES))()
        _check(self.L.cgbdsm_kernel_times(self.h, arr, 1 if reset else 0))
        return {name: {"ms": arr[i].ms, "launches": arr[i].launches, "bytes": arr[i].bytes, "items": arr[i].items}
                for i, name in enumerate(N.KERNEL_CLASSES)}

    def sync(self):
        _check(self.L.cgbdsm_sync(self.h))
