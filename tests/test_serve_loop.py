"""Serving-loop control flow on the CPU, with a fake engine that records
every call on a query handle (no GPU): a query still catching up when the
stream ends must not be used after its handle is closed."""
import threading  # noqa: F401
import time

from cgbdsm_amd.serve import Registration, Server


class FakeQuery:
    def __init__(self, eng, name):
        self.eng, self.name, self.closed = eng, name, False
        self.phase = 0

    def _use(self, what):
        if self.closed:
            self.eng.after_close.append((self.name, what))
        self.eng.calls.append((self.name, what))

    def set_layout(self, layout):
        self._use("set_layout")

    def index_init(self, keys, dv):
        self._use("index_init")

    def set_plan(self, plan):
        self._use("set_plan")

    def set_deadline(self, ms):
        self._use("set_deadline")

    def set_dfs_tuning(self, *a):
        self._use("set_dfs_tuning")

    def stage(self, kb):
        self._use("stage")
        time.sleep(self.eng.apply_s if self.name in self.eng.slow else 0)

    def transfer_update_all(self):
        self._use("transfer_update_all")

    def transfer_update_index(self):
        self._use("transfer_update_index")
        return True

    def build_local_index(self):
        self._use("build_local_index")

    def build_frontier(self):
        self._use("build_frontier")

    def batch_update_global_index(self):
        self._use("batch_update_global_index")

    def sync(self):
        self._use("sync")

    def close(self):
        self._use("close")
        self.closed = True


class FakeEngine:
    def __init__(self, slow=(), apply_s=0.05):
        self.calls, self.after_close, self.slow, self.apply_s = [], [], set(slow), apply_s
        self.lock = threading.Lock()

    def query(self, tables):
        return FakeQuery(self, tables)

    def memory(self):
        return 0, 0

    def matching_batch(self, qs):
        return [(1, None) for _ in qs]


def _reg(name, arrive_s=0.0, nbatches=4):
    return Registration(name=name, tables=name, plan=None, keys=[], dv=0, batches=[object()] * nbatches,
                        arrive_s=arrive_s, footprint=1)


def test_catching_query_is_stopped_before_its_handle_closes():
    eng = FakeEngine(slow={"late"}, apply_s=0.05)
    srv = Server(eng, autotune=False, tuning=(4, 0))
    n = 4
    # "late" arrives after the first batch and needs 50 ms per missed batch;
    # the stream ends while it is still catching up.
    arrivals = [(0.002 * b, 10) for b in range(n)]
    regs = [_reg("live", 0.0, n), _reg("late", 0.001, n)]
    res = srv.run(regs, arrivals)
    srv.close()
    time.sleep(0.3)  # a catch-up thread left running would now call into the closed handle
    assert eng.after_close == [], f"calls on closed handles: {eng.after_close}"
    closes = [c for c in eng.calls if c == ("late", "close")]
    assert len(closes) == 1
    assert len(res["counts"]["live"]) == n


class _OvfQuery:
    """Device query stand-in for DeviceServer._serve_batch: `heavy` ones
    fail their count with CGBDSM_ERR_OVERFLOW (a missed deadline)."""

    def __init__(self, name, count, heavy):
        self.name, self.count, self.heavy = name, count, heavy
        self.committed = 0

    def calig_update(self):
        pass

    def transfer_update_all(self):
        pass

    def transfer_update_index(self):
        return True

    def build_local_index(self):
        pass

    def build_frontier(self):
        pass

    def set_dfs_tuning(self, *a):
        self.tuning = a

    def set_deadline(self, ms):
        self.deadline = ms

    def profile(self, on):
        pass

    def kernel_times(self, reset=False):
        return {"dfs": {"ms": 1.0}}

    def matching(self):
        from cgbdsm_amd.engine import CgbdsmError
        if self.heavy:
            raise CgbdsmError(5, "final DFS exceeded the deadline")
        return self.count, [self.count]

    def batch_update_global_index(self):
        self.committed += 1

    def sync(self):
        pass


class _OvfEngine:
    def matching_batch(self, qs):
        from cgbdsm_amd.engine import CgbdsmError
        if any(q.heavy for q in qs):
            raise CgbdsmError(5, "final DFS exceeded the deadline")
        return [(q.count, [q.count]) for q in qs]

    def memory(self):
        return 0, 0


def test_served_batch_evicts_the_query_that_misses_its_deadline():
    """Served mode: the batched count (and the autotuner's first count) of a
    batch holding a query that misses the deadline fails; the queries are
    then counted one by one, the failing one is evicted by name and not
    committed, the others keep their counts and are committed, and the
    launch shape is tuned at the next batch without it."""
    from cgbdsm_amd.serve import DeviceServer

    srv = DeviceServer(_OvfEngine(), None, mode="served", threads=2, deadline_ms=50.0)
    live = {"a": (_OvfQuery("a", 3, False), None), "b": (_OvfQuery("b", 0, True), None),
            "c": (_OvfQuery("c", 7, False), None)}
    gone = {}
    out = srv._serve_batch(live, gone)
    assert out == {"a": 3, "c": 7}
    assert list(gone) == ["b"]
    assert live["a"][0].committed == 1 and live["c"][0].committed == 1 and live["b"][0].committed == 0
    assert srv.tuning is None
    del live["b"]
    out = srv._serve_batch(live, {})
    assert out == {"a": 3, "c": 7} and srv.tuning is not None
    # the batched launch's deadline covered both queries; restored after it
    assert live["a"][0].deadline == 50.0
    srv.close()
