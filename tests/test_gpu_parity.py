"""Device parity: the gfx950 engine against the CPU oracle, batch by batch,
through the C ABI (bit-exact index, masks, counts, per-edge split, tuples)."""
import numpy as np
import pytest

from cgbdsm_amd.engine import CgbdsmError, Engine
from cgbdsm_amd.host import HostSession, generate_files, synth_spec
from tests import oracle_lib as O
from tests import parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _log(msg):
    print(msg, flush=True)


# (name, scale, avg_degree, labels, zipf, query size, kind, seed, batch edges, ebits)
STREAMS = [
    ("cfg1-s12-tree5", 12, 8, 8, 0.0, 5, "tree", 1, None, 5),
    ("s11-dense5-b300", 11, 8, 4, 0.0, 5, "dense", 2, 300, 5),
    ("s11-sparse6-b250", 11, 8, 4, 0.0, 6, "sparse", 3, 250, 5),
    ("s12-zipf-dense6-wide", 12, 12, 6, 1.2, 6, "dense", 4, 500, 7),
    ("s10-tree7-b64", 10, 8, 3, 0.0, 7, "tree", 5, 64, 5),
    ("s11-dense7-b400", 11, 10, 3, 0.0, 7, "dense", 6, 400, 5),
]


@pytest.mark.parametrize("case", STREAMS, ids=[c[0] for c in STREAMS])
def test_stream_parity(eng, case):
    name, scale, deg, labels, zipf, qs, kind, seed, be, ebits = case
    print(f"case {name}", flush=True)
    spec = synth_spec(scale, deg, labels, zipf_s=zipf, seed=seed, query_size=qs, query_kind=kind, query_seed=seed)
    host = HostSession(spec=spec, batch_edges=be, threads=4, ebits=ebits)
    res = parity.run_stream(host, eng, log=_log)
    assert res["batches"] > 0
    assert res["mismatches"] == 0


def test_config1_full_size(eng):
    """R-MAT scale 14, degree 8, 8 uniform labels, seed 1, 90/10 split, one
    batch, 5-vertex random-walk BFS-tree query; count also against the
    closed form emb(G_1) - emb(G_0)."""
    print("case config1", flush=True)
    spec = synth_spec(14, 8, 8, seed=1, query_size=5, query_kind="tree", query_seed=1)
    host = HostSession(spec=spec, batch_edges=None, threads=8)
    offs, adj, lab = host.graph()
    ql, qe = host.query_info()
    before = O.brute_count(offs, adj, lab, ql, qe)
    res = parity.run_stream(host, eng, check_masks=True, log=_log)
    offs, adj, lab = host.graph()
    after = O.brute_count(offs, adj, lab, ql, qe)
    assert res["mismatches"] == 0
    assert res["total"] == after - before


def test_tuples_parity(eng):
    print("case tuples", flush=True)
    spec = synth_spec(9, 6, 3, seed=7, query_size=4, query_kind="dense", query_seed=7)
    host = HostSession(spec=spec, batch_edges=150, threads=4)
    res = parity.run_stream(host, eng, tuples_cap=1 << 18, log=_log)
    assert res["mismatches"] == 0


@pytest.mark.parametrize("params,dfs", [
    ({"bfs_row_limit": 0}, None),                          # final DFS straight from the seeds
    ({"bfs_row_limit": 1 << 40}, None),                    # BFS down to the last level
    ({"bfs_row_limit": 1 << 40, "frontier_cap": 32}, None),  # overflow -> DFS from the last good level
    ({"bfs_row_limit": 0, "donate_min": 2}, None),         # aggressive work donation between wavefronts
    ({"bfs_row_limit": 0, "donate_min": 0}, None),         # no donation
    ({"bfs_row_limit": 0}, {"shortest_list": False, "product_suffix": False}),  # first-list scans, full enumeration
    ({"bfs_row_limit": 0}, {"shortest_list": True, "product_suffix": False}),   # shortest list, full enumeration
    ({"bfs_row_limit": 1 << 40}, {"shortest_list": True, "product_suffix": True}),  # products from BFS rows
], ids=["dfs-only", "bfs-deep", "frontier-overflow", "steal-eager", "steal-off", "first-list", "no-product",
        "bfs-product"])
def test_matching_paths(eng, params, dfs):
    print(f"case paths {params} {dfs}", flush=True)
    spec = synth_spec(11, 8, 4, seed=8, query_size=6, query_kind="dense", query_seed=8)
    host = HostSession(spec=spec, batch_edges=300, threads=4)
    res = parity.run_stream(host, eng, match_params=params, dfs_options=dfs, check_index=False, check_masks=False,
                            log=_log)
    assert res["mismatches"] == 0


@pytest.mark.parametrize("kind,labels,seed,min_len", [("dense", 2, 21, 256), ("tree", 1, 22, 256),
                                                      ("sparse", 2, 23, 256), ("dense", 2, 24, 64),
                                                      ("sparse", 3, 25, 128), ("dense", 2, 24, 0),
                                                      ("sparse", 3, 25, 0)])
def test_wave_search_parity(eng, kind, labels, seed, min_len):
    """Shared lists searched by the whole wavefront (held one per lane up to
    64 entries, by pivots from min_len entries; min_len 0: the search turned
    off, one binary search per lane) on hubs of a scale-12, degree-16 graph
    with one to three labels: index, counts, per-edge split and tuples as
    the oracle's."""
    print(f"case wave search {kind} {min_len}", flush=True)
    spec = synth_spec(12, 16, labels, seed=seed, query_size=5, query_kind=kind, query_seed=seed)
    host = HostSession(spec=spec, batch_edges=400, threads=4)
    res = parity.run_stream(host, eng, check_masks=False, wave_search=min_len, max_batches=4, log=_log)
    assert res["batches"] > 0 and res["mismatches"] == 0


def test_duplicate_stream_pairs(eng, tmp_path):
    """Update streams that repeat pairs inside one batch: the update lists
    then hold duplicates, which the reference counts once per occurrence
    (its U lists are not uniqued); the DFS must keep to first-list scans."""
    print("case duplicate pairs", flush=True)
    spec = synth_spec(10, 8, 3, seed=12, query_size=5, query_kind="dense", query_seed=12)
    qp, dp, up = tmp_path / "q.graph", tmp_path / "d.graph", tmp_path / "u.txt"
    generate_files(spec, qp, dp, up)
    lines = up.read_text().splitlines()
    dup = []
    for i, ln in enumerate(lines):
        dup.append(ln)
        if i % 7 == 0:
            dup.append(ln)
    up.write_text("\n".join(dup) + "\n")
    host = HostSession(files=(qp, dp, up), batch_edges=200, threads=4)
    res = parity.run_stream(host, eng, log=_log)
    assert res["batches"] > 0
    assert res["mismatches"] == 0


# (name, scale, avg_degree, labels, query size, kind, seed, initial share, batch edges)
GPMA_STREAMS = [
    ("gpma-s11-dense5", 11, 8, 4, 5, "dense", 2, 0.9, 300),
    ("gpma-s12-tree5-b2000", 12, 8, 8, 5, "tree", 3, 0.9, 2000),
    # a small initial graph and large batches: windows at high levels and
    # root overflows that rebuild the array at a larger size
    ("gpma-s11-grow", 11, 10, 3, 4, "dense", 4, 0.2, 4000),
]


@pytest.mark.parametrize("case", GPMA_STREAMS, ids=[c[0] for c in GPMA_STREAMS])
def test_gpma_layout_parity(eng, case):
    """The GPMA layout yields the same lists (and counts) as the oracle's
    set algebra after every apply and every commit."""
    name, scale, deg, labels, qs, kind, seed, frac, be = case
    print(f"case {name}", flush=True)
    spec = synth_spec(scale, deg, labels, seed=seed, initial_frac=frac, query_size=qs, query_kind=kind,
                      query_seed=seed)
    host = HostSession(spec=spec, batch_edges=be, threads=4)
    res = parity.run_stream(host, eng, layout="gpma", log=_log)
    assert res["batches"] > 0
    assert res["mismatches"] == 0
    st = res["stats"]
    _log(f"gpma rebuilds {st['gpma_rebuilds']}, windows {st['gpma_windows']}, top level {st['gpma_top_level']}")
    assert st["gpma_windows"] > 0
    if name == "gpma-s11-grow":
        assert st["gpma_rebuilds"] >= 1       # the root bound was exceeded
        assert st["gpma_top_level"] >= 2      # windows above the leaf level


# Few labels: suffix positions share label classes with earlier positions
# (hits), stable intersections outgrow the LDS cache, single-list positions
# with real masks - the batched product-suffix count against the oracle.
SUFFIX_STREAMS = [
    ("sfx-s12-l2-dense5", 12, 8, 2, 5, "dense", 3, 400, None),
    ("sfx-s12-l3-dense6", 12, 10, 3, 6, "dense", 5, 400, None),
    ("sfx-s13-l2-tree5", 13, 8, 2, 5, "tree", 6, 800, None),
    ("sfx-s12-l2-sparse6", 12, 8, 2, 6, "sparse", 4, 40, 2),
    ("sfx-s12-l4-dense6", 12, 8, 4, 6, "dense", 22, 400, None),
    ("sfx-s12-l4-sparse6", 12, 8, 4, 6, "sparse", 21, 400, None),
    ("sfx-s12-l4-tree6", 12, 8, 4, 6, "tree", 20, 400, None),
    # product suffix from position 2: depth-1 frontier rows
    ("sfx-s11-l6-tree4-a", 11, 8, 6, 4, "tree", 31, 300, None),
    ("sfx-s11-l6-tree4-b", 11, 8, 6, 4, "tree", 33, 300, None),
    ("sfx-s11-l6-tree5", 11, 8, 6, 5, "tree", 35, 300, None),
    # few labels: suffix classes of two members
    ("sfx-s12-l2-tree6", 12, 8, 2, 6, "tree", 7, 400, None),
    ("sfx-s12-l3-sparse6", 12, 8, 3, 6, "sparse", 8, 400, None),
    ("sfx-s12-l3-dense5", 12, 8, 3, 5, "dense", 9, 400, None),
]


@pytest.mark.parametrize("case", SUFFIX_STREAMS, ids=[c[0] for c in SUFFIX_STREAMS])
def test_batched_suffix_parity(eng, case):
    """DFS from the depth-2 seeds (no BFS levels), so that the position
    before the product suffix is scanned by the DFS itself."""
    name, scale, deg, labels, qs, kind, seed, be, maxb = case
    print(f"case {name}", flush=True)
    spec = synth_spec(scale, deg, labels, seed=seed, query_size=qs, query_kind=kind, query_seed=seed)
    host = HostSession(spec=spec, batch_edges=be, threads=4)
    res = parity.run_stream(host, eng, check_masks=False, max_batches=maxb, match_params={"bfs_row_limit": 0},
                            log=_log)
    print(f"{name}: total {res['total']}, batched suffix steps {res['dfs_batched']}", flush=True)
    assert res["batches"] > 0
    assert res["mismatches"] == 0
    if name in ("sfx-s12-l2-dense5", "sfx-s13-l2-tree5", "sfx-s12-l4-dense6", "sfx-s12-l4-sparse6",
                "sfx-s12-l4-tree6", "sfx-s11-l6-tree4-a", "sfx-s11-l6-tree4-b", "sfx-s11-l6-tree5"):
        assert res["dfs_batched"] > 0   # every class of the suffix has one member


def test_batched_suffix_is_exercised(eng):
    """With the product suffix off the batched path is not taken; counts are
    the same either way."""
    spec = synth_spec(12, 8, 2, seed=3, query_size=5, query_kind="dense", query_seed=3)
    mp = {"bfs_row_limit": 0}
    on = parity.run_stream(HostSession(spec=spec, batch_edges=400, threads=4), eng, check_index=False,
                           check_masks=False, match_params=mp, log=_log)
    off = parity.run_stream(HostSession(spec=spec, batch_edges=400, threads=4), eng, check_index=False,
                            check_masks=False, match_params=mp, dfs_options={"product_suffix": False}, log=_log)
    assert on["mismatches"] == 0 and off["mismatches"] == 0
    assert on["counts"] == off["counts"]
    assert off["dfs_batched"] == 0
    assert on["dfs_batched"] > 0


def test_rejects_out_of_range_keys(eng):
    spec = synth_spec(8, 4, 2, seed=9, query_size=4, query_kind="tree", query_seed=9)
    host = HostSession(spec=spec, batch_edges=None, threads=2)
    keys, dv, card, avg = host.initial_keys()
    q = eng.query(host.query_tables)
    bad = np.array([(np.uint64(0) << np.uint64(59)) | (np.uint64(dv + 5) << np.uint64(32))], dtype=np.uint64)
    with pytest.raises(CgbdsmError) as ei:
        q.index_init(bad, dv)
    assert ei.value.code == 1
    q.close()


def test_call_order_is_checked(eng):
    """Out-of-order calls fail with CGBDSM_ERR_INVALID: nothing staged, and
    a new batch staged while the previous one sits between stage and commit
    (it would lose that batch's commit).  After the commit, staging works
    and the counts still equal the oracle's."""
    spec = synth_spec(8, 4, 2, seed=10, query_size=4, query_kind="tree", query_seed=10)
    host = HostSession(spec=spec, batch_edges=25, threads=2)
    keys, dv, card, avg = host.initial_keys()
    q = eng.query(host.query_tables)
    q.index_init(keys, dv)
    with pytest.raises(CgbdsmError):
        q.transfer_update_all()  # nothing staged
    plan = host.make_plan(card, avg)
    q.set_plan(plan)
    idx = O.Index(plan.ndir, host.ebits)
    idx.set_dv(dv)
    idx.init(keys)
    assert host.num_batches >= 2
    kb = host.next_batch()
    nxt = host.next_batch()
    q.stage(kb)
    with pytest.raises(CgbdsmError):
        q.stage(nxt)  # staged, not applied
    q.transfer_update_all()
    with pytest.raises(CgbdsmError):
        q.stage(nxt)  # applied, not committed
    for b in (kb, nxt):
        if b is nxt:
            q.stage(nxt)
            q.transfer_update_all()
        idx.set_dv(b.dv)
        idx.insert(b.ins)
        idx.remove(b.dele)
        idx.rebuild_update(b.upd)
        if q.transfer_update_index():
            q.build_local_index()
            cnt, per = q.matching()
            with pytest.raises(CgbdsmError):
                q.stage(nxt)  # matched, not committed
            ocnt, oper = idx.match(plan, b.edge_ok, masks=idx.masks(plan))
            assert cnt == ocnt and np.array_equal(per, oper)
            q.batch_update_global_index()
        idx.insert(b.upd)
    assert np.array_equal(q.export(0), idx.export(0))
    q.close()


def test_deadline_abort_then_recount(eng):
    """A final DFS past its deadline: the host raises the mapped abort flag,
    every wave leaves its wait and the call fails with CGBDSM_ERR_OVERFLOW.
    The same query, deadline off, then counts the same batch again (the work
    queue resynchronised past the tickets abandoned mid-launch) and the
    total and per-edge split equal the oracle's."""
    print("case deadline-abort", flush=True)
    spec = synth_spec(12, 10, 3, seed=6, query_size=7, query_kind="dense", query_seed=6)
    host = HostSession(spec=spec, batch_edges=800, threads=4)
    keys, dv, card, avg = host.initial_keys()
    q = eng.query(host.query_tables)
    q.index_init(keys, dv)
    plan = host.make_plan(card, avg)
    q.set_plan(plan)
    # every frontier row goes to the DFS, which enumerates (no closed-form
    # suffix), so the launch lasts long enough for its waves to poll the flag
    q.set_match_params(bfs_row_limit=0)
    q.set_dfs_options(shortest_list=False, product_suffix=False)
    idx = O.Index(plan.ndir, host.ebits)
    idx.set_dv(dv)
    idx.init(keys)
    kb = host.next_batch()
    q.stage(kb)
    q.transfer_update_all()
    idx.set_dv(kb.dv)
    idx.insert(kb.ins)
    idx.remove(kb.dele)
    idx.rebuild_update(kb.upd)
    assert q.transfer_update_index()
    q.build_local_index()
    q.set_deadline(0.001)
    with pytest.raises(CgbdsmError) as ex:
        q.matching()
    assert ex.value.code == 5, ex.value  # CGBDSM_ERR_OVERFLOW
    print(f"aborted: {ex.value}", flush=True)
    q.set_deadline(0)
    cnt, per = q.matching()
    ocnt, oper = idx.match(plan, kb.edge_ok, masks=idx.masks(plan), threads=4)
    assert cnt == ocnt and np.array_equal(per, oper), (cnt, ocnt, list(per), list(oper))
    assert cnt > 0
    # and once more, on a queue that has now completed normally
    cnt2, per2 = q.matching()
    assert cnt2 == ocnt and np.array_equal(per2, oper)
    q.close()
