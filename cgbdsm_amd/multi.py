"""Many registered queries over one resident data graph, with Stage 1 on the
device (cgbdsm_graph_* / cgbdsm_calig_*).

The reference runs one query per process, each with its own CaLiG and its
own copy of the graph (main_calig.cpp:29-35).  Here the graph and the update
stream live once in HBM and every query's CaLiG index, main index and
update index stay resident beside them.  Per update batch the graph applies
the batch once; then, for every query, Stage 1 (cgbdsm_calig_update) and
Stage 2 run.  Two ways to run a set of queries over a batch:

  plain   the queries one after another, each with its own final-DFS launch
          (what running the queries one per process amounts to on one GPU);
  served  Stage 1, index maintenance, masks and frontiers of all queries
          concurrently (one host thread and one stream per query), then ONE
          final-DFS launch counts every query from a shared device work
          queue (cgbdsm_matching_batch), then the commits concurrently.

Counts are identical either way (every call is exact); only the schedule
differs.
"""
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .engine import CgbdsmError

OOM = 2  # CGBDSM_ERR_OOM
from .host import Planner


@dataclass
class QuerySpec:
    name: str
    labels: np.ndarray
    edges: np.ndarray  # (qe, 2)

    @property
    def qe(self):
        return int(self.edges.shape[0])


def sample_mix(workload, n, seed, sizes=(4, 12), kinds=("tree", "sparse", "dense")):
    """n seeded random-walk queries: sizes uniform in [sizes[0], sizes[1]],
    kinds a seeded shuffle of an equal share of each kind."""
    rng = np.random.default_rng(seed)
    kind_of = [kinds[i % len(kinds)] for i in range(n)]
    rng.shuffle(kind_of)
    out = []
    for i in range(n):
        size = int(rng.integers(sizes[0], sizes[1] + 1))
        qseed = int(seed * 1000 + i)
        # A dense query needs an induced walk subgraph of average degree >= 3
        # (a 4-vertex one is a 4-clique): when none is found at the drawn
        # size, the nearest size that has one is taken (the name shows it).
        tried = sorted(range(sizes[0], sizes[1] + 1), key=lambda s: (abs(s - size), s))
        for s in tried:
            try:
                lab, ed = workload.query(s, kind_of[i], qseed)
                break
            except RuntimeError:
                continue
        else:
            raise RuntimeError(f"no {kind_of[i]} query of any size in {sizes} could be sampled")
        tag = f"{kind_of[i]}{s}" + (f"(drawn {size})" if s != size else "")
        out.append(QuerySpec(f"q{i:02d}-{tag}-s{qseed}", lab, ed))
    return out


@dataclass
class Member:
    spec: QuerySpec
    planner: Planner
    q: object  # DeviceQuery
    counts: list = field(default_factory=list)


class QuerySet:
    """One device query per spec, attached to `graph` (device Stage 1)."""

    def __init__(self, eng, graph, specs, layout="gpma", deadline_ms=60_000, threads=16, errors=None,
                 stop_bytes=0):
        """With `errors` (a dict), a query whose attach fails for device
        memory (CGBDSM_ERR_OOM), or that would start once the engine holds
        more than `stop_bytes` (0: no limit), is recorded there by name and
        left out instead of raising."""
        self.eng, self.graph = eng, graph
        self.members = []
        self.pool = ThreadPoolExecutor(max_workers=max(1, min(threads, len(specs))))
        for sp in specs:
            if errors is not None and stop_bytes and eng.memory()[0] > stop_bytes:
                errors[sp.name] = (f"OOM: device memory: the engine holds {eng.memory()[0] / 2**30:.1f} GiB, "
                                   f"over the {stop_bytes / 2**30:.1f} GiB left for this set")
                continue
            ebits = 5 if 2 * sp.qe <= 32 else 7
            pl = Planner(sp.labels, sp.edges, ebits=ebits)
            q = eng.query(pl.query_tables)
            q.set_layout(layout)
            try:
                card, avg = q.calig_attach(graph, sp.labels)
            except CgbdsmError as e:
                q.close()
                pl.close()
                if errors is None or e.code != OOM:
                    raise
                errors[sp.name] = str(e)
                continue
            q.set_plan(pl.make_plan(card, avg))
            q.set_deadline(deadline_ms)
            self.members.append(Member(sp, pl, q))

    def names(self):
        return [m.spec.name for m in self.members]

    def drop(self, names):
        keep = []
        for m in self.members:
            if m.spec.name in names:
                m.q.close()
                m.planner.close()
            else:
                keep.append(m)
        self.members = keep

    # -- one batch (the graph has applied it) ------------------------------
    @staticmethod
    def _plain_one(m):
        m.q.calig_update()
        c, _, _ = m.q.run_batch()
        return c

    def run_plain(self, errors=None):
        """The batch for every query, one after another.  With `errors` (a
        dict), a query whose count fails (deadline, 64-bit overflow) is
        recorded there by name instead of raising; its later calls are
        skipped by the caller dropping it."""
        out = []
        for m in self.members:
            try:
                out.append(self._plain_one(m))
            except CgbdsmError as e:
                if errors is None:
                    raise
                errors[m.spec.name] = str(e)
                out.append(None)
        return out

    @staticmethod
    def _prepare(m):
        m.q.calig_update()
        m.q.transfer_update_all()
        if not m.q.transfer_update_index():
            return False
        m.q.build_local_index()
        m.q.build_frontier()
        return True

    def run_served(self, tune=None, profile_lead=False):
        """The batch for every query: Stage 1 and the frontiers concurrently,
        one batched final-DFS launch, commits concurrently.  Returns the
        counts and the lead query's DFS kernel time when profiled; the wall
        time of the three phases is kept in `last_phases` (seconds:
        prepare, matching, commit)."""
        ms = self.members
        t0 = time.perf_counter()
        live = list(self.pool.map(self._prepare, ms))
        t1 = t2 = time.perf_counter()
        ks = [k for k in range(len(ms)) if live[k]]
        out = [0] * len(ms)
        dfs = None
        if ks:
            qs = [ms[k].q for k in ks]
            if tune is not None:
                tune(qs)
            lead = qs[0]
            if profile_lead:
                lead.profile(True)
                lead.kernel_times(reset=True)
            for k, (c, _) in zip(ks, self.eng.matching_batch(qs)):
                out[k] = c
            if profile_lead:
                kt = lead.kernel_times()
                lead.profile(False)
                dfs = kt["dfs"]
            t2 = time.perf_counter()
            list(self.pool.map(lambda k: ms[k].q.batch_update_global_index(), ks))
        for m in ms:
            m.q.sync()
        self.last_phases = (t1 - t0, t2 - t1, time.perf_counter() - t2)
        return out, dfs

    def close(self):
        for m in self.members:
            m.q.close()
            m.planner.close()
        self.members = []
        self.pool.shutdown()
