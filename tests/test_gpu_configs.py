"""The benchmark configurations at full size on the GPU.

config 2: R-MAT scale 20, average degree 16, 16 labels Zipf(1.2), seed 1,
90/10 split, 100 000-edge batches.  The configuration names a 7-vertex dense
random-walk query; on this graph every such query has more than 1e11 new
embeddings per batch, beyond exact enumeration in a test, so the test uses
the 6-vertex dense query of the same sampler (the test id says so).
Stage 1 runs on the device and on the host: lit bits, compressed ids and
update keys agree every batch; the GAMMA and GPMA layouts hold identical
main indexes after every apply and commit, equal to the oracle's index fed
with the host producer's keys; the count of batch 0 equals the oracle's and
every batch's count is the same with and without shortest-list scans.

config 3 (index stress): R-MAT scale 22, average degree 32, 32 uniform
labels, 80 % initial, 10 M insertions in 1 M-edge batches, 9-vertex sparse
walk query, device Stage 1: GAMMA and GPMA main and update indexes equal
after every apply and commit on all 10 batches, and equal to the oracle's
index fed with the device's own keys on the first two (the device keys
themselves are pinned by the device-vs-host Stage-1 tests at smaller
scales).  The configuration measures index maintenance; its query has more
than 9e14 new embeddings in the first batch (the final DFS had not finished
after 120 s), so no counts are taken here: counts are pinned against the
oracle at the smaller scales.

config 4 (the multi-query headline): R-MAT scale 21, average degree 24, 16
labels Zipf(1.2), seed 1, 90/10 split, 250 000-edge batches, 64 random-walk
queries of 4-12 vertices (mix seed 1), device Stage 1, GPMA layout.  Batch 0
is the probe (a query whose count is not taken within 1 s is left out, as in
the benchmark); on the next two batches the per-query plain path and the
served path (concurrent preparation, one batched final-DFS launch over a
shared work queue) give equal counts for every kept query.  The kept
queries' counts are about 1e13 per batch, so they are pinned against the
oracle only at the smaller scales (the batched launch in test_gpu_batch.py,
the serving loop in test_gpu_serve.py)."""
import numpy as np
import pytest

from cgbdsm_amd.engine import Engine
from cgbdsm_amd.host import HostSession, Planner, Workload, synth_spec
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def _edges_of(offs, adj):
    n = offs.size - 1
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(offs).astype(np.int64))
    e = np.stack([src, adj[: int(offs[-1])]], axis=1)
    return e[e[:, 0] < e[:, 1]]


def test_config2_dense6_deviation_from_dense7():
    print("case config2 (6-vertex dense instead of 7)", flush=True)
    spec = synth_spec(20, 16, 16, zipf_s=1.2, seed=1, query_size=6, query_kind="dense", query_seed=1)
    host = HostSession(spec=spec, batch_edges=100_000, threads=16)
    keys, dv, card, avg = host.initial_keys()
    offs, adj, lab = host.graph()
    ql, qe = host.query_info()
    n = lab.size
    eng = Engine(0)
    g = eng.graph(lab, _edges_of(offs, adj))
    g.set_stream(host.update_stream().astype(np.int32).reshape(-1))
    qs = {}
    for layout in ("gamma", "gpma"):
        q = eng.query(host.query_tables)
        q.set_layout(layout)
        dcard, davg = q.calig_attach(g, ql)
        assert np.array_equal(dcard, card) and np.array_equal(davg, avg)
        qs[layout] = q
    plan = host.make_plan(card, avg)
    for q in qs.values():
        q.set_plan(plan)
    idx = O.Index(plan.ndir, 5)
    idx.set_dv(dv)
    idx.init(keys)
    assert np.array_equal(qs["gamma"].export(0), idx.export(0))
    for b in range(host.num_batches):
        print(f"config2 batch {b}", flush=True)
        kb = host.next_batch()
        g.apply_range(b * 200_000, 200_000)
        for q in qs.values():
            q.calig_update()
        assert np.array_equal(qs["gamma"].calig_lit(), host.lit(n)), b
        assert np.array_equal(qs["gpma"].calig_old_ids(), host.old_ids()), b
        assert np.array_equal(np.sort(qs["gamma"].staged_keys("upd")), np.sort(kb.upd)), b
        idx.set_dv(kb.dv)
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        idx.rebuild_update(kb.upd)
        for q in qs.values():
            q.transfer_update_all()
        m_ref = idx.export(0)
        for q in qs.values():
            assert np.array_equal(q.export(0), m_ref), b
        counts = []
        for layout, q in qs.items():
            assert q.transfer_update_index()
            q.build_local_index()
            c, per = q.matching()
            if layout == "gpma":
                q.set_dfs_options(shortest_list=False)
                c2, per2 = q.matching()  # the same batch again, first-list scans
                q.set_dfs_options(shortest_list=True)
                assert c2 == c and np.array_equal(per2, per), b
            counts.append((c, per))
        assert counts[0][0] == counts[1][0] and np.array_equal(counts[0][1], counts[1][1]), b
        if b == 0:
            om = idx.masks(plan)
            oc, oper = idx.match(plan, kb.edge_ok, masks=om, threads=16)
            assert counts[0][0] == oc and np.array_equal(counts[0][1], oper)
        for q in qs.values():
            q.batch_update_global_index()
        idx.insert(kb.upd)
        m_ref = idx.export(0)
        for q in qs.values():
            assert np.array_equal(q.export(0), m_ref), b
        print(f"config2 batch {b}: {counts[0][0]} matches", flush=True)
    for q in qs.values():
        q.close()
    g.close()
    eng.close()


def test_config3_index_stress():
    print("case config3", flush=True)
    spec = synth_spec(22, 32, 32, zipf_s=0.0, seed=1, initial_frac=0.8, stream_edges=10_000_000)
    w = Workload(spec)
    ql, qe = w.query(9, "sparse", 1)
    pl = Planner(ql, qe)
    eng = Engine(0)
    g = eng.graph(w.labels, w.edges)
    g.set_stream(w.stream)
    qs = {}
    card = avg = None
    for layout in ("gamma", "gpma"):
        q = eng.query(pl.query_tables)
        q.set_layout(layout)
        c, a = q.calig_attach(g, ql)
        if card is not None:
            assert np.array_equal(c, card) and np.array_equal(a, avg)
        card, avg = c, a
        qs[layout] = q
    plan = pl.make_plan(card, avg)
    for q in qs.values():
        q.set_plan(plan)
    m0 = qs["gamma"].export(0)
    assert np.array_equal(m0, qs["gpma"].export(0))
    idx = O.Index(plan.ndir, 5)
    idx.set_dv(qs["gamma"].stats()["dv_count"])
    idx.init(m0)
    nb = (w.stream.size + 1_999_999) // 2_000_000
    assert nb == 10
    for b in range(nb):
        print(f"config3 batch {b}", flush=True)
        g.apply_range(b * 2_000_000, 2_000_000)
        staged = {}
        for layout, q in qs.items():
            q.calig_update()
            staged[layout] = {k: q.staged_keys(k) for k in ("ins", "del", "upd")}
        for k in ("ins", "del", "upd"):
            assert np.array_equal(np.sort(staged["gamma"][k]), np.sort(staged["gpma"][k])), (b, k)
        for q in qs.values():
            q.transfer_update_all()
        m_g, m_p = qs["gamma"].export(0), qs["gpma"].export(0)
        assert np.array_equal(m_g, m_p), b
        check = b < 2
        if check:
            idx.set_dv(qs["gamma"].stats()["dv_count"])
            idx.insert(staged["gamma"]["ins"])
            idx.remove(staged["gamma"]["del"])
            idx.rebuild_update(staged["gamma"]["upd"])
            assert np.array_equal(m_g, idx.export(0)), b
        for q in qs.values():
            q.transfer_update_index()
        assert np.array_equal(qs["gamma"].export(1), qs["gpma"].export(1)), b
        if check:
            assert np.array_equal(qs["gamma"].export(1), idx.export(1)), b
        for q in qs.values():
            q.batch_update_global_index()
        if check:
            idx.insert(staged["gamma"]["upd"])
        m_g = qs["gamma"].export(0)
        assert np.array_equal(m_g, qs["gpma"].export(0)), b
        if check:
            assert np.array_equal(m_g, idx.export(0)), b
        print(f"config3 batch {b}: {m_g.size} index entries, {staged['gamma']['upd'].size} update keys", flush=True)
    for q in qs.values():
        q.close()
    g.close()
    eng.close()
    pl.close()
    w.close()


def test_config4_headline_mix():
    from cgbdsm_amd.multi import QuerySet, sample_mix

    print("case config4", flush=True)
    spec = synth_spec(21, 24, 16, zipf_s=1.2, seed=1, initial_frac=0.9, stream_edges=0)
    w = Workload(spec)
    specs = sample_mix(w, 64, 1)
    budget = 240 * 2**30
    eng = Engine(0, budget_bytes=budget)
    g = eng.graph(w.labels, w.edges)
    g.set_stream(w.stream)
    errors = {}
    plain = QuerySet(eng, g, specs, layout="gpma", deadline_ms=1000, errors=errors, stop_bytes=budget // 2)
    bi = 2 * 250_000
    g.apply_range(0, bi)
    plain.run_plain(errors)
    plain.drop(set(errors))
    served = QuerySet(eng, g, [m.spec for m in plain.members], layout="gpma", deadline_ms=60_000, errors=errors,
                      stop_bytes=int(budget * 0.85))
    plain.drop(set(errors))
    assert plain.names() == served.names() and len(plain.names()) >= 4, (plain.names(), served.names())
    print(f"config4: {len(plain.names())} queries kept: {plain.names()}", flush=True)
    for m in plain.members:
        m.q.set_deadline(60_000)
    for b in (1, 2):
        g.apply_range(b * bi, bi)
        cp = plain.run_plain()
        cs, _ = served.run_served()
        assert cp == cs, (b, cp, cs)
        assert sum(cs) > 0
        print(f"config4 batch {b}: {sum(cs)} matches, plain == served", flush=True)
    plain.close()
    served.close()
    g.close()
    eng.close()
    w.close()
