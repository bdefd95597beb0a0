"""Batched matching: several registered queries over one update stream are
counted by ONE final-DFS launch (cgbdsm_matching_batch); every query's
count and per-edge split must equal the CPU oracle's, batch by batch."""
import numpy as np
import pytest

from cgbdsm_amd.engine import CgbdsmError, Engine
from cgbdsm_amd.host import HostSession, synth_spec
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

# (query kind, query size, query seed) over the same data graph and stream
MIX = [("dense", 5, 2), ("tree", 5, 3), ("sparse", 6, 4), ("dense", 4, 5), ("tree", 6, 6), ("dense", 6, 7)]


@pytest.fixture(scope="module")
def eng():
    e = Engine(0)
    yield e
    e.close()


def _open(eng, mix, scale=11, batch_edges=300, donate_min=None):
    out = []
    for kind, size, qseed in mix:
        spec = synth_spec(scale, 8, 4, seed=2, query_size=size, query_kind=kind, query_seed=qseed)
        host = HostSession(spec=spec, batch_edges=batch_edges, threads=4)
        keys, dv, card, avg = host.initial_keys()
        q = eng.query(host.query_tables)
        if donate_min is not None:
            q.set_match_params(donate_min=donate_min)
        q.index_init(keys, dv)
        plan = host.make_plan(card, avg)
        q.set_plan(plan)
        idx = O.Index(plan.ndir, host.ebits)
        idx.set_dv(dv)
        idx.init(keys)
        out.append((host, q, idx, plan))
    return out


@pytest.mark.parametrize("donate_min", [None, 1, 0])
def test_matching_batch_parity(eng, donate_min):
    print(f"case batch donate_min={donate_min}", flush=True)
    sess = _open(eng, MIX, donate_min=donate_min)
    nb = min(h.num_batches for h, _, _, _ in sess)
    assert nb >= 3
    totals = [0] * len(sess)
    for b in range(nb):
        live = []
        for k, (host, q, idx, plan) in enumerate(sess):
            kb = host.next_batch()
            q.stage(kb)
            idx.set_dv(kb.dv)
            q.transfer_update_all()
            idx.insert(kb.ins)
            idx.remove(kb.dele)
            nonempty = q.transfer_update_index()
            idx.rebuild_update(kb.upd)
            if nonempty:
                q.build_local_index()
                if (k + b) % 2 == 0:
                    q.build_frontier()   # prebuilt frontier for half of them
                live.append((k, kb))
        # vary the leader: rotate the order of the batch's queries
        live = live[b % max(1, len(live)):] + live[: b % max(1, len(live))]
        got = eng.matching_batch([sess[k][1] for k, _ in live])
        for (k, kb), (cnt, per) in zip(live, got):
            host, q, idx, plan = sess[k]
            ocnt, oper = idx.match(plan, kb.edge_ok, threads=4)
            print(f"batch {b} query {k} ({MIX[k]}): {cnt} vs oracle {ocnt}", flush=True)
            assert cnt == ocnt and np.array_equal(per, oper), (b, k, cnt, ocnt, list(per), list(oper))
            st = q.stats()
            assert st["last_matches"] == cnt
            totals[k] += cnt
        for k, kb in live:
            sess[k][1].batch_update_global_index()
            sess[k][2].insert(kb.upd)
    print(f"totals {totals}", flush=True)
    assert sum(totals) > 0
    for _, q, _, _ in sess:
        q.close()


def test_matching_batch_single_and_tuples(eng):
    """A batch of one query equals cgbdsm_matching; the embeddings of a query
    counted in a batch are available afterwards."""
    print("case batch single", flush=True)
    sess = _open(eng, [("dense", 4, 9)], scale=9, batch_edges=150)
    host, q, idx, plan = sess[0]
    for _ in range(host.num_batches):
        kb = host.next_batch()
        q.stage(kb)
        idx.set_dv(kb.dv)
        q.transfer_update_all()
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        ne = q.transfer_update_index()
        idx.rebuild_update(kb.upd)
        if ne:
            q.build_local_index()
            [(cnt, per)] = eng.matching_batch([q])
            ocnt, oper, otup, otei = idx.match(plan, kb.edge_ok, tuples_cap=1 << 16)
            assert cnt == ocnt and np.array_equal(per, oper)
            n, tup, tei = q.tuples(1 << 16)
            assert n == ocnt
            a = np.concatenate([tup, tei[:, None].astype(np.uint32)], axis=1)
            bb = np.concatenate([otup, otei[:, None].astype(np.uint32)], axis=1)
            assert np.array_equal(a[np.lexsort(a.T[::-1])], bb[np.lexsort(bb.T[::-1])])
            q.batch_update_global_index()
        idx.insert(kb.upd)
    q.close()


def test_matching_batch_rejects(eng):
    print("case batch rejects", flush=True)
    sess = _open(eng, [("tree", 4, 11), ("tree", 4, 12)], scale=8, batch_edges=None)
    (h0, q0, _, _), (h1, q1, _, _) = sess
    with pytest.raises(CgbdsmError):
        eng.matching_batch([q0])             # not after build_local_index
    for host, q, _, _ in sess:
        q.stage(host.next_batch())
        q.transfer_update_all()
        if q.transfer_update_index():
            q.build_local_index()
    with pytest.raises(CgbdsmError):
        eng.matching_batch([q0, q0])         # listed twice
    with pytest.raises(CgbdsmError):
        eng.matching_batch([])
    for _, q, _, _ in sess:
        q.close()


def test_dfs_tuning_keeps_counts(eng):
    """Every final-DFS launch shape the autotuner tries, and every stealing
    and search knob the full tuner tries, counts the same batch identically
    (checked by the tuners themselves) and equal to the oracle; a second
    cgbdsm_matching on the same batch counts it again."""
    from cgbdsm_amd.tune import autotune, autotune_full

    print("case dfs tuning", flush=True)
    sess = _open(eng, [("dense", 5, 2), ("sparse", 6, 4), ("tree", 5, 3)], scale=11, batch_edges=300)
    tuned = 0
    for b in range(2):
        live = []
        for k, (host, q, idx, plan) in enumerate(sess):
            kb = host.next_batch()
            q.stage(kb)
            idx.set_dv(kb.dv)
            q.transfer_update_all()
            idx.insert(kb.ins)
            idx.remove(kb.dele)
            ne = q.transfer_update_index()
            idx.rebuild_update(kb.upd)
            if ne:
                q.build_local_index()
                live.append((k, kb))
        if not live:
            continue
        qs = [sess[k][1] for k, _ in live]
        # one query alone (b == 0), then all of them in one launch
        sel = qs[:1] if b == 0 else qs
        if b == 0:
            best, table, got = autotune(eng, sel, grabs=(1, 4, 16, 64), blocks_per_cu=(0, 1))
            assert len(table) == 8 and best in table
        else:
            # the full tuner: every stealing and search candidate must give
            # the counts of the first, which are checked against the oracle
            cand = dict(range_min=(256, 4096), donate_min=(1, 16), cache_cap=(0, 256), wave_search=(64, 1024, 0))
            knobs, table, got = autotune_full(eng, sel, grabs=(1, 16), blocks_per_cu=(0,), **cand)
            assert len(table["shape"]) == 2
            for name, values in cand.items():
                assert set(table[name]) == set(values) and knobs[name] in values
        tuned += 1
        for (k, kb), (cnt, per) in zip(live, got):
            host, q, idx, plan = sess[k]
            ocnt, oper = idx.match(plan, kb.edge_ok, threads=4)
            assert cnt == ocnt and np.array_equal(per, oper), (b, k, cnt, ocnt)
        if b == 0:
            c2, per2 = qs[0].matching()        # recount of the same batch
            assert c2 == got[0][0] and np.array_equal(per2, got[0][1])
            for q in qs[1:]:
                q.matching()
        for k, kb in live:
            sess[k][1].batch_update_global_index()
            sess[k][2].insert(kb.upd)
    assert tuned > 0
    with pytest.raises(CgbdsmError):
        sess[0][1].set_dfs_tuning(2048, 0)
    for _, q, _, _ in sess:
        q.close()


def test_two_batched_launches_at_once(eng):
    """Two groups of queries, each counted by its own batched launch from its
    own host thread, so that one group's index maintenance and frontier
    builds overlap the other group's launch: every count equals the count of
    the same query run alone.  (A batched launch runs on its first query's
    stream and reads the other queries' frontiers, so it must wait for each
    frontier's completion event on that query's own stream.)"""
    import threading

    print("case two batched launches at once", flush=True)
    groups = [MIX[:3], MIX[3:]]
    sess = [_open(eng, g) for g in groups]
    nb = min(h.num_batches for g in sess for h, _, _, _ in g)
    assert nb >= 3
    # Stage 1 on this thread, before the device work starts
    kbs = [[[h.next_batch() for _ in range(nb)] for h, _, _, _ in g] for g in sess]
    got = [[[0] * len(g) for _ in range(nb)] for g in groups]
    errors = []

    def run(gi):
        try:
            qs = [q for _, q, _, _ in sess[gi]]
            for b in range(nb):
                live = []
                for k, q in enumerate(qs):
                    q.stage(kbs[gi][k][b])
                    q.transfer_update_all()
                    if q.transfer_update_index():
                        q.build_local_index()
                        q.build_frontier()
                        live.append(k)
                if live:
                    for k, (c, _) in zip(live, eng.matching_batch([qs[k] for k in live])):
                        got[gi][b][k] = c
                for k in live:
                    qs[k].batch_update_global_index()
                for q in qs:
                    q.sync()
        except BaseException as e:  # raised on the test thread below
            errors.append(e)

    th = [threading.Thread(target=run, args=(gi,)) for gi in range(len(groups))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for g in sess:
        for _, q, _, _ in g:
            q.close()
    assert not errors, errors[0]
    for gi, g in enumerate(groups):
        for k, (host, q, _, _) in enumerate(_open(eng, g)):
            for b in range(nb):
                q.stage(kbs[gi][k][b])
                c, _, _ = q.run_batch()
                assert c == got[gi][b][k], (gi, k, b, c, got[gi][b][k])
            q.close()
    print(f"groups {groups}: {nb} batches, counts equal alone", flush=True)
