"""Host logic of the final-DFS tuners (cgbdsm_amd/tune.py) on stand-in
queries: the fastest candidate of every knob is kept and left applied, a
candidate whose counts differ from the first is an error, and the match
parameters other than the donation threshold survive the search.  The
device side (every candidate counting a real batch equal to the oracle) is
tests/test_gpu_batch.py::test_dfs_tuning_keeps_counts."""
import numpy as np
import pytest

from cgbdsm_amd import tune


class FakeQuery:
    """Records the knobs it is given; its 'DFS time' is a function of them."""

    def __init__(self, cost, counts=(7, (1, 2, 4)), bad=None):
        self.cost, self.counts, self.bad = cost, counts, bad
        self.grab, self.bpc = 4, 0
        self.range_min, self.cache_cap, self.donate_min, self.wave = 1024, 0, 4, 64
        self.profiled, self.ms = False, 0.0

    def set_dfs_tuning(self, g, bpc):
        self.grab, self.bpc = g, bpc

    def set_dfs_steal(self, range_min, cache_cap):
        self.range_min, self.cache_cap = range_min, cache_cap

    def set_match_params(self, bfs_row_limit=65536, frontier_cap=0, donate_min=4):
        self.match_params = (bfs_row_limit, frontier_cap, donate_min)
        self.donate_min = donate_min

    def set_dfs_search(self, wave=True, min_len=64):
        self.wave = min_len if wave else 0

    def knobs(self):
        return dict(rows_per_grab=self.grab, blocks_per_cu=self.bpc, range_min=self.range_min,
                    donate_min=self.donate_min, cache_cap=self.cache_cap, wave_search=self.wave)

    def profile(self, on):
        self.profiled = on

    def kernel_times(self, reset=False):
        if reset:
            self.ms = 0.0
        return {"dfs": {"ms": self.ms}}

    def matching(self):
        k = self.knobs()
        self.ms = self.cost(k)
        c, per = self.counts
        if self.bad and self.bad(k):
            c += 1
        return c, np.array(per, dtype=np.uint64)


class FakeEngine:
    def matching_batch(self, qs):
        out = [q.matching() for q in qs]
        qs[0].ms = sum(q.ms for q in qs)  # the lead's events time the one launch
        return out


def _cost(k):
    # minimum at grab 8, 2 blocks per CU, range 4096, donate 1, cache 256, wave search 256
    return (abs(k["rows_per_grab"] - 8) + abs(k["blocks_per_cu"] - 2) + (k["range_min"] != 4096)
            + (k["donate_min"] != 1) + (k["cache_cap"] != 256) + (k["wave_search"] != 256) + 10.0)


def test_autotune_picks_and_applies_the_fastest_shape():
    qs = [FakeQuery(_cost), FakeQuery(_cost)]
    best, table, counts = tune.autotune(FakeEngine(), qs, grabs=(1, 8, 16), blocks_per_cu=(0, 2))
    assert best == (8, 2) and len(table) == 6
    assert all((q.grab, q.bpc) == (8, 2) for q in qs)
    assert counts[0][0] == 7 and not qs[0].profiled


def test_autotune_full_walks_every_knob():
    q = FakeQuery(_cost)
    q.set_match_params(1000, 77, 4)
    knobs, table, _ = tune.autotune_full(FakeEngine(), [q], grabs=(8, 16), blocks_per_cu=(2,))
    assert knobs == dict(rows_per_grab=8, blocks_per_cu=2, range_min=4096, donate_min=1, cache_cap=256,
                         wave_search=256)
    assert q.knobs() == knobs  # left applied
    assert q.match_params == (1000, 77, 1)  # row limit and frontier cap kept
    assert set(table) == {"shape", "range_min", "donate_min", "cache_cap", "wave_search"}
    assert set(table["wave_search"]) == set(tune.WAVE_SEARCH)
    assert not q.profiled


@pytest.mark.parametrize("bad", [lambda k: k["rows_per_grab"] == 16, lambda k: k["wave_search"] == 0])
def test_a_candidate_that_changes_the_counts_is_an_error(bad):
    q = FakeQuery(_cost, bad=bad)
    with pytest.raises(RuntimeError, match="changed the counts"):
        tune.autotune_full(FakeEngine(), [q], grabs=(8, 16), blocks_per_cu=(2,))
