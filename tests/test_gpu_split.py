"""CPU/GPU split of the final DFS (cgbdsm_set_cpu_split; the reference's
Step5CPUGPUSplit, graph/graph_gpu.cu:1185-1258, with the host count of
index/cpu_dfs.cpp:173-217 over the host index mirror of
index/cpu_index_mirror.h:32-48), through the C ABI.

With the host taking a fixed share of every frontier, each batch's count and
per-edge split equal the oracle's, and the host mirror's M (after apply and
after commit) and U equal the oracle's index; with the moving average on,
the share follows the measured balance and the counts stay exact; deletion
streams with Stage 1 on the device keep the mirror exact through the
negative counts' extra Stage-2 pass."""
import numpy as np
import pytest

from cgbdsm_amd.engine import Engine
from cgbdsm_amd.host import HostSession, synth_spec
from tests import oracle_lib as O
from tests import parity
from tests.parity import apply_pairs, csr_of, deletion_stream

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _log(msg):
    print(msg, flush=True)


# (name, scale, degree, labels, query size, kind, seed, batch edges, layout, host share)
FIXED = [
    ("s11-dense5-half", 11, 8, 4, 5, "dense", 2, 300, "gamma", 0.5),
    ("s11-sparse6-tenth", 11, 8, 4, 6, "sparse", 3, 250, "gpma", 0.1),
    ("s10-tree7-most", 10, 8, 3, 7, "tree", 5, 64, "gamma", 0.9),
]


@pytest.mark.parametrize("case", FIXED, ids=[c[0] for c in FIXED])
def test_fixed_share(eng, case):
    name, scale, deg, labels, qs, kind, seed, be, layout, ratio = case
    print(f"case split {name}", flush=True)
    spec = synth_spec(scale, deg, labels, seed=seed, query_size=qs, query_kind=kind, query_seed=seed)
    host = HostSession(spec=spec, batch_edges=be, threads=4)
    res = parity.run_stream(host, eng, layout=layout, log=_log,
                            split=dict(threads=4, ratio=ratio, alpha=0.0, min_rows=1))
    assert res["mismatches"] == 0
    assert res["split"], "no batch was split"
    assert sum(s["last_cpu_matches"] for s in res["split"]) > 0
    for s in res["split"]:
        assert s["ratio"] == ratio and s["last_cpu_rows"] > 0 and s["last_gpu_rows"] > 0


def test_moving_share(eng):
    """Share 0 probes with one row per host thread, then follows the
    measured balance (alpha 0.5); counts exact on every batch."""
    print("case split moving share", flush=True)
    spec = synth_spec(11, 8, 4, seed=4, query_size=5, query_kind="tree", query_seed=4)
    host = HostSession(spec=spec, batch_edges=400, threads=4)
    res = parity.run_stream(host, eng, log=_log, split=dict(threads=4, ratio=0.0, alpha=0.5, min_rows=1))
    assert res["mismatches"] == 0
    assert res["split"]
    assert 1 <= res["split"][0]["last_cpu_rows"] <= 4  # the probe: one row per host thread
    for s in res["split"]:
        assert 0.0 <= s["ratio"] <= 0.95
        assert s["last_gpu_ms"] > 0 and s["last_cpu_ms"] > 0


def test_deletions_with_device_stage1(eng):
    """Mixed insert/delete stream, Stage 1 on the device: positive and
    negative counts with the host taking half of each frontier equal the
    oracle's anchored counts, and the mirror's M equals the device's index
    after every commit."""
    print("case split deletions", flush=True)
    spec = synth_spec(11, 10, 4, seed=4, query_size=6, query_kind="dense", query_seed=4)
    host = HostSession(spec=spec, batch_edges=200, threads=4)
    host.initial_keys()
    offs, adj, lab = host.graph()
    ql, qe = host.query_info()
    n = lab.size
    lists = [adj[int(offs[v]):int(offs[v + 1])].tolist() for v in range(n)]
    cur = {(v, int(w)) for v in range(n) for w in lists[v] if v < int(w)}
    batches = deletion_stream(n, lists, host.update_stream().reshape(-1).tolist(), 200, np.random.default_rng(4), 0.5)
    g = eng.graph(lab, np.array(sorted(cur), dtype=np.uint32).reshape(-1, 2))
    g.set_stream(np.array([x for pairs in batches for ab in pairs for x in ab], dtype=np.int32))
    q = eng.query(host.query_tables)
    card, avg = q.calig_attach(g, ql)
    plan = host.make_plan(card, avg)
    q.set_plan(plan)
    q.set_cpu_split(threads=4, ratio=0.5, alpha=0.0, min_rows=1)
    start, split, host_matches = 0, 0, 0
    for b, pairs in enumerate(batches):
        o0, a0 = csr_of(n, cur)
        g.apply_range(start, 2 * len(pairs))
        start += 2 * len(pairs)
        cur, ins, dele = apply_pairs(cur, pairs, n)
        q.calig_update()
        nc, nper = q.calig_match_deleted()
        onc, onper = O.anchored_count(o0, a0, lab, ql, qe, np.array(dele, dtype=np.uint32).reshape(-1, 2))
        assert nc == onc and np.array_equal(nper, onper), (b, nc, onc)
        q.transfer_update_all()
        assert np.array_equal(q.mirror_export(0), q.export(0)), b
        if q.transfer_update_index():
            q.build_local_index()
            sb = q.split_stats()["batches"]
            c, per = q.matching()
            st = q.split_stats()
            if st["batches"] > sb:
                split += 1
                host_matches += st["last_cpu_matches"]
        else:
            c, per = 0, np.zeros(plan.qe, dtype=np.uint64)
        o1, a1 = csr_of(n, cur)
        oc, oper = O.anchored_count(o1, a1, lab, ql, qe, np.array(ins, dtype=np.uint32).reshape(-1, 2))
        assert c == oc and np.array_equal(per, oper), (b, c, oc)
        q.batch_update_global_index()
        assert np.array_equal(q.mirror_export(0), q.export(0)), b
        print(f"split deletions batch {b}: {c} matches, {nc} destroyed", flush=True)
    assert split > 0 and host_matches > 0
    q.set_cpu_split(threads=0)
    assert q.split_stats()["threads"] == 0
    q.close()
    g.close()
