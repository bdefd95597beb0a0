"""Continuous-batching serving loop over one device engine.

Requests arrive open-loop: update batches of one shared edge stream, and
query registrations (with an optional departure batch).  The loop keeps
every admitted query's index (main index, GPMA array, update lists) resident
in device memory and, per update batch:

  * index maintenance, masks and frontiers of all live queries run
    concurrently (one host thread and one HIP stream per query);
  * ONE persistent final-DFS launch counts every live query from a shared
    device work queue (cgbdsm_matching_batch), launch shape autotuned on the
    first batch;
  * the commits run concurrently.

A registration is admitted when a live slot is free and the engine's device
memory budget has room for its footprint next to the footprints of the
queries already admitted (and next to the bytes cgbdsm_ctx_memory reports
in use; a load that still exceeds the budget fails with CGBDSM_ERR_OOM and
is retried): otherwise it waits until a query leaves.  An admitted query is brought up to the stream position on
its own thread and stream while the live queries keep being served: its
index is loaded and the batches it missed are applied (index maintenance
only, no matching); it joins the live set at the next batch boundary it
has reached.  A query's count for a batch depends only on its index before
and after that batch, so the counts of a query served in the mix equal the
counts of the same query run alone (tests/test_gpu_serve.py checks this).

The reference has no serving loop (one query per process, batches read from
a file, main_calig.cpp:178-338); Stage 1 (CaLiG) stays per query on the host
and is supplied as per-batch key sets, as the reference's S1 thread
supplies them to its S2 loop.
"""
import statistics
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

from .engine import CgbdsmError

OOM = 2
OVERFLOW = 5


@dataclass
class Registration:
    """One query and its Stage-1 output: initial index keys and one key batch
    per stream batch (KeyBatch list, index = stream batch number)."""
    name: str
    tables: object
    plan: object
    keys: object
    dv: int
    batches: list
    arrive_s: float = 0.0            # arrival time, seconds after the loop starts
    leave_after: int = None          # last stream batch served (None: to the end)
    footprint: int = None            # device bytes to reserve at admission (None: estimated)

    def need(self):
        # index keys as lists, records, masks and frontier buffers: a coarse
        # bound when the caller has no measured footprint
        return self.footprint if self.footprint is not None else 64 * len(self.keys) + (32 << 20)


@dataclass
class _Query:
    reg: Registration
    dq: object = None
    applied: int = 0                 # stream batches applied to its index
    joined_at: int = None            # first batch counted in the live set
    ready: bool = False              # caught up, waiting to join
    failed: object = None
    stop: bool = False               # the loop ended: the catch-up thread returns
    thread: threading.Thread = None
    cond: threading.Condition = field(default_factory=threading.Condition)


class Server:
    def __init__(self, engine, *, mode="served", layout="gpma", deadline_ms=120000.0, autotune=True, max_live=128,
                 threads=16, tuning=None, log=None):
        """tuning: a DFS launch shape (rows_per_grab, blocks_per_cu) found
        earlier; otherwise (autotune) it is tuned on the first batch served."""
        if mode not in ("served", "plain"):
            raise ValueError("mode is 'served' or 'plain'")
        self.eng, self.mode, self.layout = engine, mode, layout
        self.deadline_ms, self.autotune, self.max_live = deadline_ms, autotune, max_live
        self.pool = ThreadPoolExecutor(max_workers=threads)
        self.log = log or (lambda m: None)
        self.lock = threading.Lock()
        self.b_next = 0              # next stream batch the live set will count
        self.tuning = tuple(tuning) if tuning else None

    # ------------------------------------------------------------ catch-up
    def _open(self, q):
        r = q.reg
        dq = self.eng.query(r.tables)
        try:
            dq.set_layout(self.layout)
            dq.index_init(r.keys, r.dv)
            dq.set_plan(r.plan)
            dq.set_deadline(self.deadline_ms)
            if self.tuning:
                dq.set_dfs_tuning(*self.tuning)
        except Exception:
            dq.close()
            raise
        q.dq = dq

    def _apply(self, q, b):
        """Index maintenance of stream batch b without matching."""
        dq = q.dq
        dq.stage(q.reg.batches[b])
        dq.transfer_update_all()
        dq.transfer_update_index()
        dq.batch_update_global_index()
        dq.sync()

    def _catch_up(self, q):
        """Load q and apply the batches it missed until it reaches the live
        set's position; returns when it joined, or when the loop sets q.stop
        (the loop then joins this thread before it closes q's handle)."""
        try:
            self._open(q)
            while True:
                with q.cond:
                    if q.stop:
                        return
                with self.lock:
                    target = self.b_next
                if q.applied < target:
                    self._apply(q, q.applied)
                    q.applied += 1
                    continue
                with q.cond:
                    q.ready = True
                    # wait until joined, stopped, or the live set moved on
                    q.cond.wait_for(lambda: q.stop or q.joined_at is not None or self.b_next > q.applied,
                                    timeout=1.0)
                    if q.joined_at is not None or q.stop:
                        return
                    q.ready = False
        except Exception as e:  # admission failure (OOM) or device error
            q.failed = e
            if q.dq is not None:
                q.dq.close()
                q.dq = None
            with q.cond:
                q.ready = True
                q.cond.notify_all()

    def _stop_catching(self, catching):
        """Stop and join the catch-up threads of queries that never joined."""
        for q in catching:
            with q.cond:
                q.stop = True
                q.cond.notify_all()
        for q in catching:
            if q.thread is not None:
                q.thread.join()

    # ------------------------------------------------------------ one batch
    def _prepare(self, q, b):
        dq = q.dq
        dq.stage(q.reg.batches[b])
        dq.transfer_update_all()
        if not dq.transfer_update_index():
            return False
        dq.build_local_index()
        dq.build_frontier()
        return True

    def _plain_one(self, q, b):
        dq = q.dq
        dq.stage(q.reg.batches[b])
        c, _, _ = dq.run_batch()
        return c

    def _serve_batch(self, live, b):
        out = {}
        if self.mode == "plain":
            for q in live:
                out[q.reg.name] = self._plain_one(q, b)
            return out
        ok = list(self.pool.map(lambda q: self._prepare(q, b), live))
        ks = [q for q, f in zip(live, ok) if f]
        for q, f in zip(live, ok):
            if not f:
                out[q.reg.name] = 0
        if ks:
            qs = [q.dq for q in ks]
            if self.autotune and self.tuning is None:
                from .tune import autotune
                best, _, _ = autotune(self.eng, qs, grabs=(2, 4, 8), blocks_per_cu=(0,))
                self.tuning = best
                for q in live:
                    q.dq.set_dfs_tuning(*best)
            for q, (c, _) in zip(ks, self.eng.matching_batch(qs)):
                out[q.reg.name] = c
            list(self.pool.map(lambda q: q.dq.batch_update_global_index(), ks))
        for q in live:
            q.dq.sync()
        return out

    # ------------------------------------------------------------ the loop
    def run(self, registrations, arrivals):
        """registrations: Registration list; arrivals: [(arrival_s, pairs)] per
        stream batch, in stream order (seconds after the loop starts, which is
        after the queries registered at time <= 0 are loaded).  Returns counts
        {name: {batch: count}}, per-batch latency (ms, arrival -> every live
        count available), and admission records."""
        pending = sorted(registrations, key=lambda r: r.arrive_s)
        waiting, catching, live = [], [], []
        counts = {r.name: {} for r in registrations}
        lat, admitted, rejected_oom = [], {}, 0
        blocked = False  # a load failed for memory: wait for a departure
        # Queries registered before the stream starts (arrive_s <= 0) are
        # loaded first, within the budget, and are live from batch 0.
        while pending and pending[0].arrive_s <= 0 and len(live) < self.max_live:
            used, budget = self.eng.memory()
            reserved = sum(q.reg.need() for q in live)
            if budget and max(used, reserved) + pending[0].need() > budget:
                break
            q = _Query(pending.pop(0))
            self._open(q)
            q.joined_at = 0
            live.append(q)
            admitted[q.reg.name] = {"arrived_s": q.reg.arrive_s, "admitted_s": 0.0, "at_batch": 0,
                                    "joined_batch": 0}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        n = len(arrivals)
        t_first = arrivals[0][0] if arrivals else 0.0
        while self.b_next < n:
            b = self.b_next
            # registrations that have arrived wait for admission
            while pending and pending[0].arrive_s <= now():
                waiting.append(pending.pop(0))
            # admission: a free slot and room in the memory budget
            used, budget = self.eng.memory()
            while waiting and not blocked and len(live) + len(catching) < self.max_live:
                # reservations: every admitted query holds its footprint
                reserved = sum(q.reg.need() for q in live + catching)
                if budget and max(used, reserved) + waiting[0].need() > budget:
                    blocked = True  # no room: wait for a departure
                    break
                r = waiting.pop(0)
                q = _Query(r)
                catching.append(q)
                admitted[r.name] = {"arrived_s": r.arrive_s, "admitted_s": now(), "at_batch": b}
                q.thread = threading.Thread(target=self._catch_up, args=(q,), daemon=True)
                q.thread.start()
            # joins at this batch boundary; failed loads go back to the queue
            for q in list(catching):
                with q.cond:
                    if not q.ready:
                        continue
                    if q.failed is not None:
                        catching.remove(q)
                        q.thread.join()
                        if isinstance(q.failed, CgbdsmError) and q.failed.code == OOM:
                            rejected_oom += 1
                            blocked = True
                            waiting.insert(0, q.reg)
                            admitted.pop(q.reg.name, None)
                            continue
                        raise q.failed
                    if q.applied == b:
                        q.joined_at = b
                        catching.remove(q)
                        live.append(q)
                        admitted[q.reg.name]["joined_batch"] = b
                    q.cond.notify_all()
            # the batch itself
            t_arr = arrivals[b][0]
            if now() < t_arr:
                time.sleep(min(t_arr - now(), 0.001))
                continue
            res = self._serve_batch(live, b) if live else {}
            t_done = now()
            for name, c in res.items():
                counts[name][b] = c
            lat.append(1e3 * (t_done - t_arr))
            with self.lock:
                self.b_next = b + 1
            for q in catching:
                with q.cond:
                    q.cond.notify_all()
            # departures after this batch free their memory
            for q in [q for q in live if q.reg.leave_after is not None and q.reg.leave_after <= b]:
                live.remove(q)
                q.dq.close()
                q.dq = None
                blocked = False
        t_end = now()
        # Queries still catching up stop applying batches before their
        # handles close: a thread still queuing work on a closed handle
        # would touch freed device buffers.
        self._stop_catching(catching)
        for q in live + catching:
            if q.dq is not None:
                q.dq.close()
                q.dq = None
        edges = sum(p for _, p in arrivals)
        return {
            "counts": counts,
            "latency_ms": lat,
            "admissions": admitted,
            "oom_retries": rejected_oom,
            "update_edges_per_s": edges / max(t_end - t_first, 1e-9),
            "p50_ms": statistics.median(lat) if lat else 0.0,
            "p99_ms": sorted(lat)[min(len(lat) - 1, int(0.99 * len(lat)))] if lat else 0.0,
            "dfs_tuning": self.tuning,
        }

    def close(self):
        self.pool.shutdown()


@dataclass
class QueryRequest:
    """A query registration for the device-Stage-1 server: the query itself
    (labels, edges); its CaLiG index is built on the device over the
    resident graph when it is admitted."""
    name: str
    labels: object
    edges: object
    arrive_s: float = 0.0
    leave_after: int = None
    footprint: int = None            # device bytes to reserve (None: measured at the first admission)


class DeviceServer:
    """Serving loop with Stage 1 on the device.

    The data graph and the whole update stream stay resident in HBM
    (cgbdsm_graph_*); every admitted query keeps its CaLiG index, main index
    and update lists resident beside them.  Update batches arrive open-loop
    (arrival time, size in stream edges) and are applied in stream order;
    query registrations arrive at their own times.  Per batch:

      * the graph applies the batch once;
      * served: every live query's Stage 1 (cgbdsm_calig_update), index
        maintenance, masks and frontier run concurrently (one host thread and
        one stream per query), ONE persistent final-DFS launch counts them
        all, and the commits run concurrently;
        plain: the live queries one after another, each with its own launch;
      * meanwhile, queries admitted at this batch boundary build their CaLiG
        index over the graph as it now stands (cgbdsm_calig_attach, on their
        own threads and streams, overlapping the live set's work) and join
        the live set at the next batch.

    Admission: a free slot (max_live) and room in the engine's memory budget
    for the query's footprint next to the admitted queries' footprints; a
    load that still fails with CGBDSM_ERR_OOM waits for a departure.  A
    query's count for a batch depends only on its index before and after the
    batch, so its counts in the mix equal its counts alone."""

    def __init__(self, engine, graph, *, mode="served", layout="gpma", deadline_ms=120000.0, max_live=128,
                 threads=16, tuning=None, log=None):
        if mode not in ("served", "plain"):
            raise ValueError("mode is 'served' or 'plain'")
        self.eng, self.graph, self.mode, self.layout = engine, graph, mode, layout
        self.deadline_ms, self.max_live = deadline_ms, max_live
        self.pool = ThreadPoolExecutor(max_workers=threads)
        self.log = log or (lambda m: None)
        self.tuning = tuple(tuning) if tuning else None
        self.footprint = None

    def _attach(self, r):
        from .host import Planner
        ebits = 5 if 2 * len(r.edges) <= 32 else 7
        pl = Planner(r.labels, r.edges, ebits=ebits)
        dq = self.eng.query(pl.query_tables)
        try:
            used0, _ = self.eng.memory()
            dq.set_layout(self.layout)
            card, avg = dq.calig_attach(self.graph, r.labels)
            dq.set_plan(pl.make_plan(card, avg))
            dq.set_deadline(self.deadline_ms)
            if self.tuning:
                dq.set_dfs_tuning(*self.tuning)
            used1, _ = self.eng.memory()
            r.measured = used1 - used0
        except Exception:
            dq.close()
            pl.close()
            raise
        return dq, pl

    def _prepare(self, dq):
        dq.calig_update()
        dq.transfer_update_all()
        if not dq.transfer_update_index():
            return False
        dq.build_local_index()
        dq.build_frontier()
        return True

    def _plain_one(self, dq):
        dq.calig_update()
        c, _, _ = dq.run_batch()
        return c

    def _serve_batch(self, live, evicted):
        """Counts of every live query for the batch the graph just applied.
        A query whose count cannot be taken (deadline, or a count beyond 64
        bits: CGBDSM_ERR_OVERFLOW) is evicted and recorded in `evicted`;
        in served mode the batch's queries are then counted one by one to
        find it."""
        out = {}
        names = list(live)
        qs = [live[n][0] for n in names]
        if self.mode == "plain":
            for n, dq in zip(names, qs):
                try:
                    out[n] = self._plain_one(dq)
                except CgbdsmError as e:
                    if e.code != OVERFLOW:
                        raise
                    evicted[n] = str(e)[:200]
            return out
        ok = list(self.pool.map(self._prepare, qs))
        ks = [k for k, f in enumerate(ok) if f]
        for k, f in enumerate(ok):
            if not f:
                out[names[k]] = 0
        if ks:
            # The batched launch (led by its first query) counts every query's
            # batch: its deadline is the per-query deadline times the queries.
            lead = qs[ks[0]]
            lead.set_deadline(self.deadline_ms * len(ks))
            try:
                # The launch shape is tuned on the first batch whose batched
                # count succeeds (the autotuner counts the batch once per shape).
                if self.tuning is None:
                    from .tune import autotune
                    best, _, _ = autotune(self.eng, [qs[k] for k in ks], grabs=(2, 4, 8), blocks_per_cu=(0,))
                    self.tuning = best
                    for dq in qs:
                        dq.set_dfs_tuning(*best)
                for k, (c, _) in zip(ks, self.eng.matching_batch([qs[k] for k in ks])):
                    out[names[k]] = c
            except CgbdsmError as e:
                if e.code != OVERFLOW:
                    raise
                lead.set_deadline(self.deadline_ms)
                for k in ks:
                    try:
                        out[names[k]] = qs[k].matching()[0]
                    except CgbdsmError as e2:
                        if e2.code != OVERFLOW:
                            raise
                        evicted[names[k]] = str(e2)[:200]
            lead.set_deadline(self.deadline_ms)
            list(self.pool.map(lambda k: qs[k].batch_update_global_index(),
                               [k for k in ks if names[k] not in evicted]))
        for dq in qs:
            dq.sync()
        return out

    def run(self, requests, arrivals):
        """requests: QueryRequest list; arrivals: [(arrival_s, edges)] per
        update batch in stream order (seconds after the loop starts, which is
        after the queries registered at time <= 0 are attached).  Returns
        counts {name: {batch: count}}, per-batch latency (ms, arrival -> every
        live count available) and admission records."""
        pending = sorted(requests, key=lambda r: r.arrive_s)
        waiting, live = [], {}
        counts = {r.name: {} for r in requests}
        lat, admitted, oom, evictions = [], {}, 0, {}
        blocked = False

        def fits(r):
            used, budget = self.eng.memory()
            need = r.footprint if r.footprint is not None else (self.footprint or 0)
            return not budget or used + need <= budget

        while pending and pending[0].arrive_s <= 0 and len(live) < self.max_live and fits(pending[0]):
            r = pending.pop(0)
            live[r.name] = self._attach(r)
            self.footprint = max(self.footprint or 0, r.measured)
            admitted[r.name] = {"arrived_s": r.arrive_s, "admitted_s": 0.0, "joined_batch": 0}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        t_first = arrivals[0][0] if arrivals else 0.0
        start = 0
        for b, (t_arr, edges) in enumerate(arrivals):
            while now() < t_arr:
                time.sleep(min(t_arr - now(), 0.001))
            while pending and pending[0].arrive_s <= now():
                waiting.append(pending.pop(0))
            self.graph.apply_range(start, 2 * edges)
            start += 2 * edges
            # admissions at this boundary attach over the graph after batch b
            joining = []
            while waiting and not blocked and len(live) + len(joining) < self.max_live and fits(waiting[0]):
                r = waiting.pop(0)
                joining.append((r, self.pool.submit(self._attach, r)))
                admitted[r.name] = {"arrived_s": r.arrive_s, "admitted_s": now(), "joined_batch": b + 1}
            gone = {}
            res = self._serve_batch(live, gone) if live else {}
            t_done = now()
            for name, c in res.items():
                counts[name][b] = c
            lat.append(1e3 * (t_done - t_arr))
            self.log(f"batch {b}: {edges} edges, {lat[-1]:.1f} ms after arrival, {len(res)} live counts"
                     + (f", evicting {sorted(gone)}" if gone else ""))
            for name, why in gone.items():
                dq, pl = live.pop(name)
                dq.close()
                pl.close()
                evictions[name] = {"batch": b, "reason": why}
                blocked = False
            for r, fut in joining:
                try:
                    live[r.name] = fut.result()
                    self.footprint = max(self.footprint or 0, r.measured)
                except CgbdsmError as e:
                    if e.code != OOM:
                        raise
                    oom += 1
                    blocked = True
                    waiting.insert(0, r)
                    admitted.pop(r.name, None)
            for name in [n for n, (dq, pl) in live.items() if self._leaves(requests, n, b)]:
                dq, pl = live.pop(name)
                dq.close()
                pl.close()
                blocked = False
        t_end = now()
        for dq, pl in live.values():
            dq.close()
            pl.close()
        edges = sum(p for _, p in arrivals)
        return {
            "counts": counts, "latency_ms": lat, "admissions": admitted, "oom_retries": oom,
            "update_edges_per_s": edges / max(t_end - t_first, 1e-9),
            "p50_ms": statistics.median(lat) if lat else 0.0,
            "p99_ms": sorted(lat)[min(len(lat) - 1, int(0.99 * len(lat)))] if lat else 0.0,
            "dfs_tuning": self.tuning, "footprint_bytes": self.footprint, "evicted": evictions,
        }

    @staticmethod
    def _leaves(requests, name, b):
        for r in requests:
            if r.name == name:
                return r.leave_after is not None and r.leave_after <= b
        return False

    def close(self):
        self.pool.shutdown()
