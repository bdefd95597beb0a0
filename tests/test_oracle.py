"""Pins of the CPU oracle (runs without a GPU).

The reference ships no tests, fixtures or golden outputs, so every oracle
function is pinned against mathematics that fixes its answer:
  * brute force vs textbook closed forms (triangles, 2-paths, k-stars, cliques)
  * anchored per-edge counts vs the closed form emb(G_b) - emb(G_{b-1})
  * index algebra vs numpy set operations
  * masks: pure prune (count with masks == count without) and a direct
    recomputation of the definition
  * full CPU path (host Stage 1 + oracle Stage 2) vs the closed form
A deliberately planted error in the oracle's inputs must be caught.
"""
import itertools
import math

import numpy as np
import pytest

from cgbdsm_amd.host import HostSession, synth_spec
from tests import oracle_lib as O


def csr(n, edges):
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    offs = np.zeros(n + 1, dtype=np.uint64)
    flat = []
    for v in range(n):
        offs[v + 1] = offs[v] + len(adj[v])
        flat.extend(sorted(adj[v]))
    return offs, np.array(flat if flat else [0], dtype=np.uint32)


def random_graph(n, m, seed):
    rng = np.random.default_rng(seed)
    es = set()
    while len(es) < m:
        a, b = rng.integers(0, n, 2)
        if a != b:
            es.add((min(a, b), max(a, b)))
    return sorted(es)


def degrees(offs):
    return np.diff(offs).astype(np.int64)


# ---------------------------------------------------------------- brute force

def test_brute_triangle_equals_six_times_triangles():
    n = 40
    edges = random_graph(n, 200, 1)
    offs, adj = csr(n, edges)
    lab = np.zeros(n, dtype=np.int32)
    es = set(edges)
    tri = sum(1 for a, b, c in itertools.combinations(range(n), 3)
              if (a, b) in es and (b, c) in es and (a, c) in es)
    got = O.brute_count(offs, adj, lab, [0, 0, 0], [0, 1, 1, 2, 0, 2])
    assert got == 6 * tri


def test_brute_two_path_and_star_closed_forms():
    n = 50
    offs, adj = csr(n, random_graph(n, 160, 2))
    lab = np.zeros(n, dtype=np.int32)
    d = degrees(offs)
    assert O.brute_count(offs, adj, lab, [0, 0, 0], [0, 1, 1, 2]) == int(np.sum(d * (d - 1)))
    for k in (2, 3, 4):
        q_edges = []
        for i in range(1, k + 1):
            q_edges += [0, i]
        want = sum(math.perm(int(x), k) for x in d)
        assert O.brute_count(offs, adj, lab, [0] * (k + 1), q_edges) == want


def test_brute_clique_counts_on_complete_graph():
    n = 9
    offs, adj = csr(n, [(a, b) for a in range(n) for b in range(a + 1, n)])
    lab = np.zeros(n, dtype=np.int32)
    q4 = [a for e in itertools.combinations(range(4), 2) for a in e]
    assert O.brute_count(offs, adj, lab, [0] * 4, q4) == math.perm(n, 4)


def test_brute_respects_labels():
    # labelled path a-b-c where only label pattern 1-2-1 is allowed
    n = 30
    offs, adj = csr(n, random_graph(n, 90, 3))
    rng = np.random.default_rng(3)
    lab = rng.integers(0, 3, n).astype(np.int32)
    want = 0
    for v in range(n):
        if lab[v] != 2:
            continue
        nb = adj[int(offs[v]):int(offs[v + 1])]
        ones = [x for x in nb if lab[x] == 1]
        want += len(ones) * (len(ones) - 1)
    assert O.brute_count(offs, adj, lab, [1, 2, 1], [0, 1, 1, 2]) == want


# ---------------------------------------------------------------- anchored

@pytest.mark.parametrize("seed", [1, 2, 3])
def test_anchored_equals_closed_form(seed):
    n = 60
    all_edges = random_graph(n, 300, seed)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(all_edges))
    old = [all_edges[i] for i in perm[:240]]
    new = [all_edges[i] for i in perm[240:]]
    lab = rng.integers(0, 2, n).astype(np.int32)
    qlab = [0, 1, 0, 1]
    qe = [0, 1, 1, 2, 2, 3, 0, 2]  # triangle + pendant
    o0, a0 = csr(n, old)
    o1, a1 = csr(n, old + new)
    before = O.brute_count(o0, a0, lab, qlab, qe)
    after = O.brute_count(o1, a1, lab, qlab, qe)
    tot, per = O.anchored_count(o1, a1, lab, qlab, qe, np.array(new))
    assert tot == after - before
    assert per.sum() == tot


def test_anchored_catches_planted_error():
    """Planting one wrong 'new' edge changes the anchored total."""
    n = 40
    edges = random_graph(n, 150, 7)
    old, new = edges[:120], edges[120:]
    lab = np.zeros(n, dtype=np.int32)
    o1, a1 = csr(n, edges)
    qe = [0, 1, 1, 2, 0, 2]
    good, _ = O.anchored_count(o1, a1, lab, [0, 0, 0], qe, np.array(new))
    bad, _ = O.anchored_count(o1, a1, lab, [0, 0, 0], qe, np.array(new[:-1] + [old[0]]))
    o0, a0 = csr(n, old)
    assert good == O.brute_count(o1, a1, lab, [0, 0, 0], qe) - O.brute_count(o0, a0, lab, [0, 0, 0], qe)
    assert bad != good


# ---------------------------------------------------------------- index algebra

def _keys(rng, n, ndir, dv, ebits=5):
    e = rng.integers(0, ndir, n).astype(np.uint64)
    u = rng.integers(0, dv, n).astype(np.uint64)
    v = rng.integers(0, dv, n).astype(np.uint64)
    return (e << np.uint64(64 - ebits)) | (u << np.uint64(32)) | v


@pytest.mark.parametrize("ebits", [5, 7])
def test_index_algebra_matches_numpy_sets(ebits):
    rng = np.random.default_rng(11)
    ndir, dv = 6, 300
    idx = O.Index(ndir, ebits)
    idx.set_dv(dv)
    a = _keys(rng, 3000, ndir, dv, ebits)
    idx.init(a)
    ref = np.unique(a)
    assert np.array_equal(idx.export(0), ref)
    for step in range(4):
        ins = _keys(rng, 800, ndir, dv, ebits)
        dele = np.concatenate([_keys(rng, 300, ndir, dv, ebits), rng.choice(ref, 200)])
        idx.insert(ins)
        idx.remove(dele)
        ref = np.setdiff1d(np.union1d(ref, ins), dele)
        assert np.array_equal(idx.export(0), ref), step
        upd = np.concatenate([_keys(rng, 100, ndir, dv, ebits)] * 2)  # duplicates are kept in U
        idx.rebuild_update(upd)
        assert np.array_equal(idx.export(1), np.sort(upd))


# ---------------------------------------------------------------- full CPU path

def split_mismatch(per, ref):
    """Query-edge tags whose counts differ between two per-edge splits."""
    per, ref = np.asarray(per, dtype=np.uint64), np.asarray(ref, dtype=np.uint64)
    return [e for e in range(max(per.size, ref.size)) if e >= per.size or e >= ref.size or per[e] != ref[e]]


def match_rows(plan, tup, tei, old_ids):
    """or_match tuples (compressed ids in global-order positions) as rows of
    original ids in query-vertex order, with the tag as the last column."""
    rows = np.zeros((len(tup), plan.qv + 1), dtype=np.int64)
    for p in range(plan.qv):
        rows[:, plan.go_vs[p]] = old_ids[tup[:, p]] if len(tup) else 0
    rows[:, plan.qv] = tei
    return rows


def rows_of(tup, tags):
    return np.concatenate([np.asarray(tup, dtype=np.int64).reshape(len(tags), -1),
                           np.asarray(tags, dtype=np.int64)[:, None]], axis=1)


def rows_equal(a, b):
    if a.shape != b.shape:
        return False
    return np.array_equal(a[np.lexsort(a.T[::-1])], b[np.lexsort(b.T[::-1])])


def _cpu_stream(spec, batch_edges, check_li=False, threads=4, tuples_cap=0):
    host = HostSession(spec=spec, batch_edges=batch_edges, threads=threads)
    keys, dv, card, avg = host.initial_keys()
    plan = host.make_plan(card, avg)
    ql, qe = host.query_info()
    stream = host.update_stream()
    pos = 0
    idx = O.Index(plan.ndir, 5)
    idx.set_dv(dv)
    idx.init(keys)
    offs, adj, lab = host.graph()
    prev = O.brute_count(offs, adj, lab, ql, qe)
    if check_li:
        li = host.lit(len(lab))
        assert np.array_equal(li, O.li_fixpoint(offs, adj, lab, ql, qe)), "initial lit bits differ from the fixpoint"
    splits = []
    for _ in range(host.num_batches):
        kb = host.next_batch()
        pairs = stream[pos:pos + kb.pairs]
        pos += kb.pairs
        idx.set_dv(kb.dv)
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        idx.rebuild_update(kb.upd)
        masks = idx.masks(plan)
        if tuples_cap:
            cnt, per, tup, tei = idx.match(plan, kb.edge_ok, masks=masks, tuples_cap=tuples_cap)
        else:
            cnt, per = idx.match(plan, kb.edge_ok, masks=masks)
        cnt0, per0 = idx.match(plan, kb.edge_ok, masks=None)
        idx.insert(kb.upd)
        offs, adj, lab = host.graph()
        cur = O.brute_count(offs, adj, lab, ql, qe)
        assert cnt == cnt0 and np.array_equal(per, per0), "support masks must be a pure prune"
        assert cnt == cur - prev, "count differs from emb(G_b) - emb(G_b-1)"
        atot, aper = O.anchored_count(offs, adj, lab, ql, qe, pairs)
        assert atot == cnt and not split_mismatch(per, aper), \
            f"per-edge split {list(per)} differs from the anchored enumeration {list(aper)}"
        if tuples_cap:
            assert cnt <= tuples_cap
            btup, btag = O.brute_tuples(offs, adj, lab, ql, qe, pairs, cap=tuples_cap)
            assert rows_equal(match_rows(plan, tup, tei, host.old_ids()), rows_of(btup, btag)), \
                "oracle tuples differ from the brute-force new embeddings"
        if check_li:
            li = host.lit(len(lab))
            fp = O.li_fixpoint(offs, adj, lab, ql, qe)
            assert np.array_equal(li, fp), "incremental lit bits differ from the from-scratch fixpoint"
        splits.append(per)
        prev = cur
    return host, splits


@pytest.mark.parametrize("kind,seed", [("tree", 1), ("dense", 2), ("sparse", 3), ("dense", 4)])
def test_cpu_path_matches_closed_form(kind, seed):
    spec = synth_spec(10, 8, 4, seed=seed, query_size=5, query_kind=kind, query_seed=seed)
    _cpu_stream(spec, 300)


@pytest.mark.parametrize("kind,seed,scale,labels", [("tree", 7, 8, 4), ("dense", 8, 8, 3), ("sparse", 9, 8, 3),
                                                    ("dense", 10, 9, 6)])
def test_oracle_tuples_equal_brute_force(kind, seed, scale, labels):
    """The embeddings or_match enumerates (mapped to original ids and query
    vertices) and their edge tags equal the brute-force new embeddings of
    every batch; the per-edge split equals the anchored enumeration."""
    spec = synth_spec(scale, 6, labels, seed=seed, query_size=5, query_kind=kind, query_seed=seed)
    _cpu_stream(spec, 60, tuples_cap=200000)


def _one_batch(seed=8):
    spec = synth_spec(8, 6, 3, seed=seed, query_size=5, query_kind="dense", query_seed=seed)
    host = HostSession(spec=spec, batch_edges=60, threads=2)
    keys, dv, card, avg = host.initial_keys()
    plan = host.make_plan(card, avg)
    ql, qe = host.query_info()
    idx = O.Index(plan.ndir, 5)
    idx.set_dv(dv)
    idx.init(keys)
    pairs = host.update_stream()
    for _ in range(host.num_batches):
        kb = host.next_batch()
        b_pairs, pairs = pairs[:kb.pairs], pairs[kb.pairs:]
        idx.set_dv(kb.dv)
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        idx.rebuild_update(kb.upd)
        cnt, per, tup, tei = idx.match(plan, kb.edge_ok, masks=idx.masks(plan), tuples_cap=200000)
        idx.insert(kb.upd)
        if np.count_nonzero(per) >= 2 and cnt > 10:
            offs, adj, lab = host.graph()
            return plan, per, (offs, adj, lab, ql, qe, b_pairs), match_rows(plan, tup, tei, host.old_ids())
    raise AssertionError("no batch with two non-zero edge tags")


def test_split_check_catches_planted_permutation():
    """A per-edge split with two tags' counts swapped must fail the anchored
    comparison that passes on the true split."""
    plan, per, g, _ = _one_batch()
    _, aper = O.anchored_count(*g)
    assert not split_mismatch(per, aper)
    nz = [e for e in range(len(per)) if per[e]]
    bad = per.copy()
    other = next(e for e in range(len(per)) if per[e] != per[nz[0]])
    bad[nz[0]], bad[other] = per[other], per[nz[0]]
    assert split_mismatch(bad, aper)


def test_tuple_check_catches_planted_column_swap():
    """Oracle tuples with two query vertices' columns swapped (or one tag
    changed) must fail the brute-force comparison that passes on the true
    tuples."""
    plan, per, g, rows = _one_batch()
    btup, btag = O.brute_tuples(*g)
    ref = rows_of(btup, btag)
    assert rows_equal(rows, ref)
    a, b = 0, 1
    while np.array_equal(rows[:, a], rows[:, b]):
        b += 1
    swapped = rows.copy()
    swapped[:, [a, b]] = rows[:, [b, a]]
    assert not rows_equal(swapped, ref)
    retag = rows.copy()
    retag[0, plan.qv] = (retag[0, plan.qv] + 1) % plan.qe
    assert not rows_equal(retag, ref)


@pytest.mark.parametrize("scale,deg,labels,seed,batch,threads", [
    (9, 8, 3, 5, 200, 4),
    # Two label groups of query vertex 1 lose candidates of vertex 552 in one
    # delete round: a check flag per (vertex, query vertex) that skipped the
    # second group's injectivity check left (1, 552) lit after the initial
    # filter and made later batches remove index entries (nondeterministically
    # under several threads).
    (11, 6, 3, 2, 400, 1),
    (11, 6, 3, 2, 400, 4),
])
def test_lit_bits_equal_from_scratch_fixpoint(scale, deg, labels, seed, batch, threads):
    spec = synth_spec(scale, deg, labels, seed=seed, query_size=5 if scale == 9 else 6, query_kind="dense",
                      query_seed=seed)
    _cpu_stream(spec, batch, check_li=True, threads=threads)


def test_masks_match_definition():
    """Recompute one depth>=2 tile from the definition in Python."""
    spec = synth_spec(9, 8, 3, seed=6, query_size=5, query_kind="dense", query_seed=6)
    host = HostSession(spec=spec, batch_edges=300, threads=2)
    keys, dv, card, avg = host.initial_keys()
    plan = host.make_plan(card, avg)
    idx = O.Index(plan.ndir, 5)
    idx.set_dv(dv)
    idx.init(keys)
    kb = host.next_batch()
    idx.set_dv(kb.dv)
    idx.insert(kb.ins)
    idx.remove(kb.dele)
    idx.rebuild_update(kb.upd)
    m, ones, W = idx.masks(plan)
    M = idx.export(0)
    U = idx.export(1)

    def lists(keys_, e, v):
        sel = keys_[(keys_ >> np.uint64(59)) == e]
        sel = sel[((sel >> np.uint64(32)) & np.uint64((1 << 27) - 1)) == v]
        return [int(x) & 0xFFFFFFFF for x in sel]

    def bit(ei, qv, v):
        t = ei * 12 + qv
        return bool(ones[t]) or bool((int(m[t * W + v // 64]) >> (v % 64)) & 1)

    checked = 0
    for ei in range(plan.qe):
        vs = list(plan.io_vs[ei])[: plan.qv]
        for d in range(2, plan.qv):
            u0 = vs[d]
            if not plan.rebuild_v[ei][u0]:
                continue
            for v in range(kb.dv):
                ok = True
                for k in range(plan.io_bni_offs[ei][d], plan.io_bni_offs[ei][d + 1]):
                    u1 = vs[plan.io_bni[ei][k]]
                    bwd = plan.eidx[u0 * 12 + u1]
                    cand = lists(M, bwd, v)
                    if plan.dir_to_edge[bwd] <= ei:
                        cand += lists(U, bwd, v)
                    ok = ok and any(bit(ei, u1, x) for x in cand)
                assert ok == bit(ei, u0, v), (ei, d, v)
                checked += 1
    assert checked > 0


@pytest.mark.parametrize("kind,seed", [("dense", 21), ("sparse", 22), ("tree", 23)])
def test_candidate_sets_after_every_batch(kind, seed):
    """The host producer's post-batch CaLiG candidate sets cand[ui][uj] of
    every lit pair (index/calig.h:28-33) are exactly the lit neighbours of
    the query neighbour's label: the definition the device Stage 1 and the
    index keys build from.  Checked for every lit pair after every batch."""
    spec = synth_spec(9, 6, 3, seed=seed, query_size=5, query_kind=kind, query_seed=seed)
    host = HostSession(spec=spec, batch_edges=150, threads=2)
    host.initial_keys()
    ql, qe = host.query_info()
    qn = [set() for _ in ql]
    for a, b in qe.reshape(-1, 2):
        qn[a].add(int(b))
        qn[b].add(int(a))
    checked = 0
    for b in range(host.num_batches + 1):
        if b:
            host.next_batch()
        offs, adj, lab = host.graph()
        li = host.lit(lab.size)
        for v in range(lab.size):
            for ui in range(len(ql)):
                if not (li[v] >> ui) & 1:
                    continue
                nb = adj[int(offs[v]):int(offs[v + 1])]
                for uj in qn[ui]:
                    want = sorted(int(x) for x in nb if lab[x] == ql[uj] and (li[x] >> uj) & 1)
                    got = host.cand(v, ui, uj)
                    assert got is not None and got.tolist() == want, (b, v, ui, uj)
                    checked += 1
    assert checked > 1000
