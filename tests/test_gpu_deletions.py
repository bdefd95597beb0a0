"""Deletion streams through the device Stage 1 and Stage 2 (pairs of
negative ids -a-1, -b-1 delete edge (a, b), index/calig.cpp:980-989), mixed
with insertions in every batch, through the C ABI.  After every batch:

  * the lit bits equal the oracle's from-scratch greatest fixpoint of the
    current graph (pairs lit before the batch die when their support goes);
  * the main index after the commit equals the index built from scratch
    over the current graph and lit set, in the device's compressed ids (so
    the entries of deleted edges and of dead pairs are gone, and nothing
    else is);
  * the count and its per-edge split equal the oracle's anchored count of
    the embeddings of the current graph that use an inserted edge;
  * the negative count (cgbdsm_calig_match_deleted) and its split equal the
    oracle's anchored count of the embeddings of the graph before the
    batch that use a deleted edge.

The reference's GPU stage only runs insertion streams: its ConstructUpdate
(index/calig.cpp:776-804) indexes vertices with the raw stream ids, so a
negative id is out of range there."""
import numpy as np
import pytest

from cgbdsm_amd.engine import Engine
from cgbdsm_amd.host import HostSession, Planner, Workload, synth_spec
from tests import oracle_lib as O
from tests.parity import apply_pairs, csr_of, deletion_stream, scratch_index

pytestmark = pytest.mark.gpu

# (name, scale, degree, labels, query size, kind, seed, insertions per batch, deletion fraction, layout)
STREAMS = [
    ("s9-dense5", 9, 6, 3, 5, "dense", 1, 40, 0.4, "gamma"),
    ("s9-sparse6", 9, 6, 2, 6, "sparse", 2, 30, 0.4, "gpma"),
    ("s9-tree5-heavy", 9, 8, 2, 5, "tree", 3, 50, 1.0, "gamma"),
    ("s11-dense6", 11, 10, 4, 6, "dense", 4, 200, 0.5, "gpma"),
    ("s12-tree6-heavy", 12, 12, 4, 6, "tree", 5, 400, 1.0, "gamma"),
]


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("case", STREAMS, ids=[c[0] for c in STREAMS])
def test_deletion_stream(eng, case):
    name, scale, deg, labels, qs, kind, seed, be, frac, layout = case
    print(f"case deletions {name}", flush=True)
    spec = synth_spec(scale, deg, labels, seed=seed, query_size=qs, query_kind=kind, query_seed=seed)
    host = HostSession(spec=spec, batch_edges=be, threads=4)
    host.initial_keys()
    offs, adj, lab = host.graph()
    ql, qe = host.query_info()
    n = lab.size
    lists = [adj[int(offs[v]):int(offs[v + 1])].tolist() for v in range(n)]
    cur = {(v, int(w)) for v in range(n) for w in lists[v] if v < int(w)}
    rng = np.random.default_rng(seed)
    batches = deletion_stream(n, lists, host.update_stream().reshape(-1).tolist(), be, rng, frac)
    flat = np.array([x for pairs in batches for ab in pairs for x in ab], dtype=np.int32)
    g = eng.graph(lab, np.array(sorted(cur), dtype=np.uint32).reshape(-1, 2))
    g.set_stream(flat)
    q = eng.query(host.query_tables)
    q.set_layout(layout)
    card, avg = q.calig_attach(g, ql)
    plan = host.make_plan(card, avg)
    q.set_plan(plan)
    start, killed, total, negative = 0, 0, 0, 0
    for b, pairs in enumerate(batches):
        lit0 = q.calig_lit()
        o0, a0 = csr_of(n, cur)
        g.apply_range(start, 2 * len(pairs))
        start += 2 * len(pairs)
        cur, ins, dele = apply_pairs(cur, pairs, n)
        q.calig_update()
        # Embeddings of the graph before the batch that the deletions destroy.
        nc, nper = q.calig_match_deleted()
        onc, onper = O.anchored_count(o0, a0, lab, ql, qe, np.array(dele, dtype=np.uint32).reshape(-1, 2))
        assert nc == onc and np.array_equal(nper, onper), (b, nc, onc)
        negative += nc
        o1, a1 = csr_of(n, cur)
        lit = q.calig_lit()
        assert np.array_equal(lit, O.li_fixpoint(o1, a1, lab, ql, qe)), b
        killed += int(np.count_nonzero(lit0 & ~lit))
        q.transfer_update_all()
        if q.transfer_update_index():
            q.build_local_index()
            c, per = q.matching()
        else:
            c, per = 0, np.zeros(plan.qe, dtype=np.uint64)
        oc, oper = O.anchored_count(o1, a1, lab, ql, qe, np.array(ins, dtype=np.uint32).reshape(-1, 2))
        assert c == oc and np.array_equal(per, oper), (b, c, oc)
        total += c
        q.batch_update_global_index()
        want = scratch_index(o1, a1, lab, ql, qe, lit, q.calig_old_ids(), plan)
        assert np.array_equal(q.export(0), want), b
        print(f"{name} batch {b}: +{len(ins)} -{len(dele)} edges, {c} matches, {nc} destroyed, "
              f"{want.size} index entries", flush=True)
    assert killed > 0  # pairs lit before a batch died in it
    assert total > 0 and negative > 0
    q.close()
    g.close()


def test_deletion_stream_batched_launch(eng):
    """Queries on one mixed stream over the unfiltered graph, counted
    together by one batched final-DFS launch (cgbdsm_matching_batch) after
    their negative counts: every count and per-edge split equals the
    oracle's anchored counts.  One query's label pattern leaves its initial
    lit set nearly empty on this graph (an index that starts empty)."""
    print("case deletions batched", flush=True)
    w = Workload(synth_spec(11, 10, 4, seed=6))
    n = w.labels.size
    lists = [[] for _ in range(n)]
    for a, b in w.edges.tolist():
        lists[a].append(b)
        lists[b].append(a)
    cur = {(min(a, b), max(a, b)) for a, b in w.edges.tolist() if a != b}
    batches = deletion_stream(n, lists, w.stream.reshape(-1).tolist(), 200, np.random.default_rng(6), 0.6)
    g = eng.graph(w.labels, np.array(sorted(cur), dtype=np.uint32).reshape(-1, 2))
    g.set_stream(np.array([x for pairs in batches for ab in pairs for x in ab], dtype=np.int32))
    qs = []
    for size, kind, seed in ((5, "tree", 7), (5, "sparse", 8), (4, "dense", 9)):
        ql, qe = w.query(size, kind, seed)
        pl = Planner(ql, qe)
        q = eng.query(pl.query_tables)
        card, avg = q.calig_attach(g, ql)
        q.set_plan(pl.make_plan(card, avg))
        qs.append((q, ql, qe.reshape(-1), pl))
    start, total = 0, 0
    for b, pairs in enumerate(batches):
        o0, a0 = csr_of(n, cur)
        g.apply_range(start, 2 * len(pairs))
        start += 2 * len(pairs)
        cur, ins, dele = apply_pairs(cur, pairs, n)
        o1, a1 = csr_of(n, cur)
        live = []
        for q, ql, qe, _ in qs:
            q.calig_update()
            nc, nper = q.calig_match_deleted()
            onc, onper = O.anchored_count(o0, a0, w.labels, ql, qe, np.array(dele, dtype=np.uint32).reshape(-1, 2))
            assert nc == onc and np.array_equal(nper, onper), (b, nc, onc)
            q.transfer_update_all()
            if q.transfer_update_index():
                q.build_local_index()
                live.append((q, ql, qe))
        got = eng.matching_batch([q for q, _, _ in live]) if live else []
        for (q, ql, qe), (c, per) in zip(live, got):
            oc, oper = O.anchored_count(o1, a1, w.labels, ql, qe, np.array(ins, dtype=np.uint32).reshape(-1, 2))
            assert c == oc and np.array_equal(per, oper), (b, c, oc)
            total += c
        for q, _, _, _ in qs:
            q.batch_update_global_index()
        print(f"batched batch {b}: +{len(ins)} -{len(dele)} edges, {len(live)} queries counted", flush=True)
    assert total > 0
    for q, _, _, pl in qs:
        q.close()
        pl.close()
    g.close()
    w.close()


def test_empty_initial_index(eng):
    """A query with a label the graph does not have: the static filter
    leaves nothing lit, the initial index is empty (its one scanned
    capacity must read as zero, whatever the reused device memory held),
    and every batch counts 0 in both directions."""
    print("case empty initial index", flush=True)
    w = Workload(synth_spec(9, 6, 3, seed=11))
    ql, qe = w.query(4, "tree", 11)
    ql = ql.copy()
    ql[0] = 7  # no vertex has label 7
    pl = Planner(ql, qe)
    g = eng.graph(w.labels, w.edges)
    g.set_stream(w.stream)
    q = eng.query(pl.query_tables)
    card, avg = q.calig_attach(g, ql)
    assert not card.any()
    q.set_plan(pl.make_plan(card, avg))
    for b in range(2):
        g.apply_range(b * 100, 100)
        q.calig_update()
        assert q.calig_match_deleted()[0] == 0
        assert q.run_batch()[0] == 0
        assert q.export(0).size == 0
    q.close()
    pl.close()
    g.close()
    w.close()
