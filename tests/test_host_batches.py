"""Stage-1 batches of any size (cgh_next_batch_edges): the same stream cut
into the same sizes gives the same key sets as the fixed-size API, and
uneven cuts consume the whole stream in order."""
import numpy as np

from cgbdsm_amd.host import HostSession, synth_spec


def _spec():
    return synth_spec(9, 6, 3, seed=5, query_size=4, query_kind="dense", query_seed=5)


def test_sized_batches_equal_fixed_batches():
    fixed = HostSession(spec=_spec(), batch_edges=120, threads=2)
    sized = HostSession(spec=_spec(), threads=2)
    fixed.initial_keys()
    sized.initial_keys()
    total = sized.stream_edges_left()
    assert total > 0
    for _ in range(fixed.num_batches):
        a = fixed.next_batch()
        b = sized.next_batch(120)
        assert a.pairs == b.pairs and a.dv == b.dv
        for x, y in ((a.ins, b.ins), (a.dele, b.dele), (a.upd, b.upd), (a.edge_ok, b.edge_ok)):
            assert np.array_equal(x, y)
    assert sized.stream_edges_left() == 0


def test_uneven_batches_cover_the_stream():
    h = HostSession(spec=_spec(), threads=2)
    h.initial_keys()
    total = h.stream_edges_left()
    sizes, got = [7, 300, 1, 55, 10_000], []
    for k in sizes:
        if h.stream_edges_left() == 0:
            break
        got.append(h.next_batch(k).pairs)
    expect, rem = [], total
    for k in sizes:
        if rem == 0:
            break
        expect.append(min(k, rem))
        rem -= expect[-1]
    assert got == expect and sum(got) == total
