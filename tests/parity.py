"""Stream-level parity harness: drives the host pipeline batch by batch and
feeds the same keys to the device engine and to the CPU oracle (test
infrastructure; the product never imports this module).

Per batch it compares, bit for bit:
  * the main index M after TransferUpdateAll and after BatchUpdateGlobalIndex
  * the update index U
  * every support-mask tile of every edge with batch lists
  * the match count and the per-edge-tag split
  * optionally the sorted embedding tuples
"""
import numpy as np

from tests import oracle_lib as O


def _masks_equal(q, idx_masks, plan, dv, edge_ok):
    m, ones, W = idx_masks
    bad = []
    for ei in range(plan.qe):
        if not edge_ok[ei]:
            continue
        for v in range(plan.qv):
            dev_ones, words = q.mask(ei, v, dv)
            t = ei * 12 + v
            if dev_ones != bool(ones[t]):
                bad.append((ei, v, "all_ones"))
                continue
            if not dev_ones:
                ref = m[t * W:(t + 1) * W]
                if not np.array_equal(words, ref[: words.size]):
                    bad.append((ei, v, "bits"))
    return bad


def run_stream(host, eng, *, check_index=True, check_masks=True, tuples_cap=0, match_params=None,
               dfs_options=None, layout=None, max_batches=None, log=None):
    keys, dv, card, avg = host.initial_keys()
    q = eng.query(host.query_tables)
    if layout:
        q.set_layout(layout)
    if match_params:
        q.set_match_params(**match_params)
    if dfs_options:
        q.set_dfs_options(**dfs_options)
    dcard, davg = q.index_init(keys, dv)
    assert np.array_equal(dcard, card), (dcard, card)
    assert np.array_equal(davg, avg), (davg, avg)
    plan = host.make_plan(card, avg)
    q.set_plan(plan)
    idx = O.Index(plan.ndir, host.ebits)
    idx.set_dv(dv)
    idx.init(keys)
    if check_index:
        assert np.array_equal(q.export(0), idx.export(0)), "initial M differs"
    res = {"batches": 0, "total": 0, "mismatches": 0, "counts": [], "per_edge": [], "skipped": 0, "dfs_batched": 0}
    nb = host.num_batches if max_batches is None else min(max_batches, host.num_batches)
    for _ in range(nb):
        kb = host.next_batch()
        q.stage(kb)
        idx.set_dv(kb.dv)
        q.transfer_update_all()
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        if check_index:
            if not np.array_equal(q.export(0), idx.export(0)):
                res["mismatches"] += 1
                if log:
                    log(f"batch {kb.index}: M after apply differs")
        nonempty = q.transfer_update_index()
        idx.rebuild_update(kb.upd)
        res["batches"] += 1
        if not nonempty:
            assert kb.upd.size == 0
            res["skipped"] += 1
            res["counts"].append(0)
            continue
        if check_index and not np.array_equal(q.export(1), idx.export(1)):
            res["mismatches"] += 1
            if log:
                log(f"batch {kb.index}: U differs")
        q.build_local_index()
        om = idx.masks(plan)
        if check_masks:
            bad = _masks_equal(q, om, plan, kb.dv, kb.edge_ok)
            if bad:
                res["mismatches"] += 1
                if log:
                    log(f"batch {kb.index}: masks differ {bad[:5]}")
        cnt, per = q.matching()
        res["dfs_batched"] += q.stats()["last_dfs_batched"]
        if tuples_cap:
            ocnt, oper, otup, otei = idx.match(plan, kb.edge_ok, masks=om, tuples_cap=tuples_cap)
            n, tup, tei = q.tuples(tuples_cap)
            a = np.concatenate([tup, tei[:, None].astype(np.uint32)], axis=1) if n else np.zeros((0, plan.qv + 1))
            b = np.concatenate([otup, otei[:, None].astype(np.uint32)], axis=1) if ocnt else np.zeros((0, plan.qv + 1))
            a = a[np.lexsort(a.T[::-1])] if len(a) else a
            b = b[np.lexsort(b.T[::-1])] if len(b) else b
            if n != ocnt or not np.array_equal(a, b):
                res["mismatches"] += 1
                if log:
                    log(f"batch {kb.index}: tuples differ ({n} vs {ocnt})")
        else:
            ocnt, oper = idx.match(plan, kb.edge_ok, masks=om, threads=4)
        if cnt != ocnt or not np.array_equal(per, oper):
            res["mismatches"] += 1
            if log:
                log(f"batch {kb.index}: count {cnt} vs oracle {ocnt}; per-edge {list(per)} vs {list(oper)}")
        q.batch_update_global_index()
        idx.insert(kb.upd)
        if check_index and not np.array_equal(q.export(0), idx.export(0)):
            res["mismatches"] += 1
            if log:
                log(f"batch {kb.index}: M after commit differs")
        res["total"] += cnt
        res["counts"].append(cnt)
        res["per_edge"].append(per)
    res["stats"] = q.stats()
    q.close()
    return res
