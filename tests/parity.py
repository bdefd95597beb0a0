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
               dfs_options=None, layout=None, max_batches=None, log=None, wave_search=None, split=None):
    """Host-Stage-1 stream through the device, checked against the oracle
    batch by batch.  split: keyword arguments of DeviceQuery.set_cpu_split;
    the host mirror's M and U are then also compared with the oracle's
    after every apply and commit."""
    keys, dv, card, avg = host.initial_keys()
    q = eng.query(host.query_tables)
    if layout:
        q.set_layout(layout)
    if match_params:
        q.set_match_params(**match_params)
    if dfs_options:
        q.set_dfs_options(**dfs_options)
    if wave_search is not None:  # None: the library default (on, 64)
        q.set_dfs_search(bool(wave_search), 256 if wave_search is True else int(wave_search or 64))
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
    res = {"batches": 0, "total": 0, "mismatches": 0, "counts": [], "per_edge": [], "skipped": 0, "dfs_batched": 0,
           "split": []}
    if split is not None:
        q.set_cpu_split(**split)

    def mirror_check(which, what, kb):
        if split is not None and not np.array_equal(q.mirror_export(which), idx.export(which)):
            res["mismatches"] += 1
            if log:
                log(f"batch {kb.index}: host mirror {what} differs")
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
        mirror_check(0, "M after apply", kb)
        nonempty = q.transfer_update_index()
        idx.rebuild_update(kb.upd)
        mirror_check(1, "U", kb)
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
        sb = q.split_stats()["batches"] if split is not None else 0
        cnt, per = q.matching()
        res["dfs_batched"] += q.stats()["last_dfs_batched"]
        if split is not None:
            st = q.split_stats()
            if st["batches"] > sb:
                res["split"].append(st)
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
        mirror_check(0, "M after commit", kb)
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


def run_stream_s1(host, eng, batch_ints, *, layout=None, check_masks=True, check_li=False, max_batches=None,
                  log=None):
    """The same stream with Stage 1 on the device (cgbdsm_calig_*): per batch
    the device lit bits, compressed ids, staged update keys and edge_ok
    equal the host producer's, and the device index, masks and counts equal
    the oracle fed with the host producer's keys.  check_li also compares
    the lit bits with the oracle's from-scratch fixpoint on the current
    graph."""
    keys, dv, card, avg = host.initial_keys()
    offs, adj, lab = host.graph()
    ql, qe = host.query_info()
    n = lab.size
    edges = np.stack([np.repeat(np.arange(n, dtype=np.uint32), np.diff(offs).astype(np.int64)),
                      adj[: int(offs[-1])]], axis=1)
    edges = edges[edges[:, 0] < edges[:, 1]]
    g = eng.graph(lab, edges)
    stream = host.update_stream().astype(np.int32).reshape(-1)
    g.set_stream(stream)
    q = eng.query(host.query_tables)
    if layout:
        q.set_layout(layout)
    dcard, davg = q.calig_attach(g, ql)
    res = {"batches": 0, "total": 0, "mismatches": 0, "counts": [], "s1": []}

    def bad(msg):
        res["mismatches"] += 1
        if log:
            log(msg)

    if not np.array_equal(dcard, card) or not np.array_equal(davg, avg):
        bad(f"initial card/avg differ: {list(dcard)} vs {list(card)}")
    if not np.array_equal(q.calig_lit(), host.lit(n)):
        bad("initial lit bits differ from the host producer's")
    if not np.array_equal(q.calig_old_ids(), host.old_ids()):
        bad("initial compressed ids differ")
    plan = host.make_plan(card, avg)
    q.set_plan(plan)
    idx = O.Index(plan.ndir, host.ebits)
    idx.set_dv(dv)
    idx.init(keys)
    if not np.array_equal(q.export(0), idx.export(0)):
        bad("initial M differs")
    nb = host.num_batches if max_batches is None else min(max_batches, host.num_batches)
    for b in range(nb):
        kb = host.next_batch()
        g.apply_range(b * batch_ints, batch_ints)
        q.calig_update()
        res["s1"].append(q.calig_stats())
        li = q.calig_lit()
        if not np.array_equal(li, host.lit(n)):
            bad(f"batch {b}: lit bits differ from the host producer's")
        if check_li:
            o2, a2, l2 = host.graph()
            if not np.array_equal(li, O.li_fixpoint(o2, a2, l2, ql, qe)):
                bad(f"batch {b}: lit bits differ from the from-scratch fixpoint")
        if not np.array_equal(q.calig_old_ids(), host.old_ids()):
            bad(f"batch {b}: compressed ids differ")
        if not np.array_equal(np.sort(q.staged_keys("upd")), np.sort(kb.upd)):
            bad(f"batch {b}: update keys differ")
        if not np.array_equal(np.unique(q.staged_keys("del")), np.unique(kb.upd)):
            bad(f"batch {b}: delete keys are not the batch's update entries")
        st = q.stats()
        idx.set_dv(kb.dv)
        q.transfer_update_all()
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        if q.stats()["dv_count"] != kb.dv:
            bad(f"batch {b}: dv {q.stats()['dv_count']} vs {kb.dv}")
        if not np.array_equal(q.export(0), idx.export(0)):
            bad(f"batch {b}: M after apply differs")
        nonempty = q.transfer_update_index()
        idx.rebuild_update(kb.upd)
        res["batches"] += 1
        if not nonempty:
            res["counts"].append(0)
            q.batch_update_global_index()
            idx.insert(kb.upd)
            continue
        if not np.array_equal(q.export(1), idx.export(1)):
            bad(f"batch {b}: U differs")
        q.build_local_index()
        om = idx.masks(plan)
        if check_masks:
            mb = _masks_equal(q, om, plan, kb.dv, kb.edge_ok)
            if mb:
                bad(f"batch {b}: masks differ {mb[:5]}")
        cnt, per = q.matching()
        ocnt, oper = idx.match(plan, kb.edge_ok, masks=om, threads=4)
        if cnt != ocnt or not np.array_equal(per, oper):
            bad(f"batch {b}: count {cnt} vs oracle {ocnt}")
        q.batch_update_global_index()
        idx.insert(kb.upd)
        if not np.array_equal(q.export(0), idx.export(0)):
            bad(f"batch {b}: M after commit differs")
        res["total"] += cnt
        res["counts"].append(cnt)
        del st
    q.close()
    g.close()
    return res


def deletion_stream(n, adj, stream, be, rng, frac=0.4):
    """Batches of `be` insertions from the flat `stream` plus deletions (as
    pairs of negative ids -a-1, -b-1) of about frac * be edges present
    before the batch (earlier insertions among them), two of them listed
    twice, and three random pairs (mostly absent edges), shuffled together.
    `adj` is the initial adjacency (lists per vertex)."""
    cur = {(min(a, b), max(a, b)) for a in range(n) for b in adj[a]}
    out = []
    for s0 in range(0, len(stream) - 1, 2 * be):
        batch = [(stream[i], stream[i + 1]) for i in range(s0, min(s0 + 2 * be, len(stream) - 1), 2)]
        present = sorted(cur)
        k = min(len(present), int(frac * be))
        dels = [present[i] for i in rng.choice(len(present), size=k, replace=False)] if k else []
        dels += dels[:2]
        dels += [(int(rng.integers(n)), int(rng.integers(n))) for _ in range(3)]
        pairs = batch + [(-a - 1, -b - 1) for a, b in dels]
        pairs = [pairs[i] for i in rng.permutation(len(pairs))]
        gone = {(min(-a - 1, -b - 1), max(-a - 1, -b - 1)) for a, b in pairs if a < 0 and b < 0}
        added = {(min(a, b), max(a, b)) for a, b in pairs if a >= 0 and b >= 0 and a != b}
        cur = (cur - (gone & cur)) | (added - cur)
        out.append(pairs)
    return out


def apply_pairs(cur, pairs, n):
    """Edge set after one batch of `pairs` (both kinds judged against `cur`,
    the set before it); returns (new set, inserted edges, deleted edges)."""
    ins, dele = set(), set()
    for a, b in pairs:
        if a < 0 and b < 0:
            x, y = -a - 1, -b - 1
            e = (min(x, y), max(x, y))
            if x != y and x < n and y < n and e in cur:
                dele.add(e)
        elif a >= 0 and b >= 0 and a != b and a < n and b < n:
            e = (min(a, b), max(a, b))
            if e not in cur:
                ins.add(e)
    return (cur - dele) | ins, sorted(ins), sorted(dele)


def csr_of(n, edges):
    """(offs u64, adj u32) of an undirected edge set."""
    lists = [[] for _ in range(n)]
    for a, b in edges:
        lists[a].append(b)
        lists[b].append(a)
    offs = np.zeros(n + 1, dtype=np.uint64)
    adj = []
    for v in range(n):
        lists[v].sort()
        adj.extend(lists[v])
        offs[v + 1] = offs[v] + len(lists[v])
    return offs, np.array(adj if adj else [0], dtype=np.uint32)


def scratch_index(offs, adj, lab, ql, qe, lit, old_ids, plan):
    """The main index built from scratch over a graph and a lit set: the
    key (dir(ui, uj) << (64 - ebits)) | (id(v) << 32) | id(w) of every edge
    v-w and lit pairs (v, ui), (w, uj) with ui-uj a query edge, ids the
    compressed ids (old_ids[id] = original vertex); sorted."""
    new_of = {int(v): i for i, v in enumerate(old_ids)}
    qv = len(ql)
    qn = [[] for _ in range(qv)]
    for a, b in np.asarray(qe).reshape(-1, 2):
        qn[int(a)].append(int(b))
        qn[int(b)].append(int(a))
    ushift = 64 - plan.ebits
    keys = []
    for v in range(lab.size):
        lv = int(lit[v])
        if not lv:
            continue
        for w in adj[int(offs[v]):int(offs[v + 1])]:
            lw = int(lit[int(w)])
            for ui in range(qv):
                if not (lv >> ui) & 1:
                    continue
                for uj in qn[ui]:
                    if (lw >> uj) & 1:
                        e = plan.eidx[ui * 12 + uj]
                        keys.append((e << ushift) | (new_of[v] << 32) | new_of[int(w)])
    return np.array(sorted(keys), dtype=np.uint64)
