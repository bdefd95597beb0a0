"""Device Stage 1 (cgbdsm_graph_* / cgbdsm_calig_*) against the host CaLiG
producer and the oracle, batch by batch, through the C ABI: lit bits (also
against the oracle's from-scratch greatest fixpoint), compressed ids, the
staged update keys, and the Stage-2 index, masks and counts that follow."""
import pytest

from cgbdsm_amd.engine import CgbdsmError, Engine
from cgbdsm_amd.host import HostSession, synth_spec
from tests import parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _log(msg):
    print(msg, flush=True)


# (name, scale, degree, labels, zipf, query size, kind, seed, batch edges, ebits, layout, check_li)
STREAMS = [
    ("s10-tree5-b300", 10, 8, 4, 0.0, 5, "tree", 1, 300, 5, None, True),
    ("s10-dense5-b200", 10, 8, 3, 0.0, 5, "dense", 2, 200, 5, None, True),
    ("s10-sparse6-b250", 10, 8, 4, 0.0, 6, "sparse", 3, 250, 5, None, True),
    ("s11-dense6-zipf-wide", 11, 12, 6, 1.2, 6, "dense", 4, 500, 7, None, True),
    ("s12-tree7-gpma", 12, 10, 3, 0.0, 7, "tree", 5, 800, 5, "gpma", False),
    ("s12-dense7-b600", 12, 10, 2, 0.0, 7, "dense", 6, 600, 5, None, False),
    ("s13-sparse8-zipf", 13, 16, 8, 1.2, 8, "sparse", 7, 2000, 5, None, False),
]


@pytest.mark.parametrize("case", STREAMS, ids=[c[0] for c in STREAMS])
def test_device_stage1_stream(eng, case):
    name, scale, deg, labels, zipf, qs, kind, seed, be, ebits, layout, check_li = case
    print(f"case {name}", flush=True)
    spec = synth_spec(scale, deg, labels, zipf_s=zipf, seed=seed, query_size=qs, query_kind=kind, query_seed=seed)
    host = HostSession(spec=spec, batch_edges=be, threads=4, ebits=ebits)
    res = parity.run_stream_s1(host, eng, 2 * be, layout=layout, check_li=check_li, log=_log)
    assert res["batches"] > 1
    assert res["mismatches"] == 0
    assert sum(s["new_pairs"] for s in res["s1"]) > 0


def test_device_stage1_config1(eng):
    """Config 1 (R-MAT s14, degree 8, 8 uniform labels, seed 1, one batch of
    the 10 % stream, 5-vertex BFS-tree query) with Stage 1 on the device."""
    print("case s1 config1", flush=True)
    spec = synth_spec(14, 8, 8, seed=1, query_size=5, query_kind="tree", query_seed=1)
    host = HostSession(spec=spec, batch_edges=None, threads=8)
    n_ints = host.update_stream().size
    res = parity.run_stream_s1(host, eng, n_ints, check_li=True, log=_log)
    assert res["batches"] == 1 and res["mismatches"] == 0


def test_device_stage1_odd_batch_ints(eng):
    """An odd batch size in ints pairs the second id of one stream edge with
    the first of the next in later batches (main_calig.cpp:220-222); the
    device slices the resident stream the same way."""
    print("case s1 odd batch", flush=True)
    spec = synth_spec(10, 8, 3, seed=8, query_size=4, query_kind="dense", query_seed=8)
    host = HostSession(spec=spec, batch_ints=301, threads=4)
    res = parity.run_stream_s1(host, eng, 301, log=_log)
    assert res["batches"] > 1 and res["mismatches"] == 0


def test_device_stage1_call_order(eng):
    print("case s1 call order", flush=True)
    spec = synth_spec(9, 6, 3, seed=9, query_size=4, query_kind="tree", query_seed=9)
    host = HostSession(spec=spec, batch_edges=100, threads=2)
    host.initial_keys()
    offs, adj, lab = host.graph()
    import numpy as np
    n = lab.size
    edges = np.stack([np.repeat(np.arange(n, dtype=np.uint32), np.diff(offs).astype(np.int64)),
                      adj[: int(offs[-1])]], axis=1)
    g = eng.graph(lab, edges[edges[:, 0] < edges[:, 1]])
    g.set_stream(host.update_stream().astype(np.int32).reshape(-1))
    q = eng.query(host.query_tables)
    ql, _ = host.query_info()
    card, avg = q.calig_attach(g, ql)
    with pytest.raises(CgbdsmError):  # plan first
        q.calig_update()
    q.set_plan(host.make_plan(card, avg))
    with pytest.raises(CgbdsmError):  # no graph batch yet
        q.calig_update()
    g.apply_range(0, 200)
    with pytest.raises(CgbdsmError):  # the query has not consumed batch 0
        g.apply_range(200, 200)
    q.calig_update()
    with pytest.raises(CgbdsmError):  # host staging is refused for a device-Stage-1 query
        q.stage(host.next_batch())
    with pytest.raises(CgbdsmError):  # graph still has an attached query
        g.close()
    q.close()
    g.close()
