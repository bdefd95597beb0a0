#!/usr/bin/env python3
"""Stage-2 hot-path benchmark on one MI355X.

Workload: labelled R-MAT graph (a,b,c,d = .57,.19,.19,.05), scale 20,
average degree 16, 16 vertex labels drawn Zipf(1.2), seed 1; a seeded
shuffle keeps 90% of the edges as the initial graph and streams the rest in
batches of 100,000 edges; one dense query from a seeded random walk (induced
subgraph, average query degree >= 3), 6 vertices by default (7-vertex dense
random-walk queries on this graph have more than 1e11 new embeddings per
batch, beyond exact enumeration).

A step is one batch through the device hot path: TransferUpdateAll ->
TransferUpdateIndex -> BuildLocalIndexBitAll -> MatchingBitAll ->
BatchUpdateGlobalIndex.  Stage 1 (CaLiG on the host) runs for every batch
before timing and its keys are copied to HBM before the timed region
(cgbdsm_batch_enqueue), so `value` is device throughput with resident inputs.

The JSON line also carries:
  roofline      the dominant kernel class: algorithmic bytes per launch over
                its mean launch time (HIP events on the query's stream)
  cpu_baseline  the CPU oracle (restated reference CPU path) timed on a
                bounded sample of the same batches, on this host
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "update edges/s + p99 batch latency, served multi-query vs plain; exact counts"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return 0.0
    k = (len(xs) - 1) * p / 100.0
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def cpu_baseline(keys, dv, plan, batches, budget_s, threads):
    """Oracle Stage 2 (set-semantics index algebra, masks, seeding + DFS) on
    the first batches of the stream until `budget_s` seconds of CPU work."""
    from tests import oracle_lib as O

    idx = O.Index(plan.ndir, 5)
    idx.set_dv(dv)
    idx.init(keys)
    edges, secs, counts = 0, 0.0, []
    for kb in batches:
        t0 = time.perf_counter()
        idx.set_dv(kb.dv)
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        idx.rebuild_update(kb.upd)
        if kb.upd.size:
            masks = idx.masks(plan)
            cnt, _ = idx.match(plan, kb.edge_ok, masks=masks, threads=threads)
            idx.insert(kb.upd)
        else:
            cnt = 0
        secs += time.perf_counter() - t0
        edges += kb.pairs
        counts.append(int(cnt))
        if secs >= budget_s:
            break
    return edges, secs, counts


def cpu_s2_sampled(keys, dv, plan, kb, threads, target_s, log=None):
    """Oracle Stage 2 time of one batch, estimated from a uniform sample of
    its update keys: the sample grows 4x at a time until its masks + count
    take `target_s` seconds (or it is the whole batch); the estimate is the
    sample's time x (update keys / sampled keys).  Index maintenance of the
    batch (insert, remove) is timed in full.  Returns (seconds, fraction
    sampled, count of the sample)."""
    import numpy as np

    from tests import oracle_lib as O

    idx = O.Index(plan.ndir, 5)
    idx.set_dv(dv)
    idx.init(keys)
    t0 = time.perf_counter()
    idx.set_dv(kb.dv)
    idx.insert(kb.ins)
    idx.remove(kb.dele)
    t_idx = time.perf_counter() - t0
    total = int(kb.upd.size)
    if not total:
        return t_idx, 1.0, 0
    perm = np.random.default_rng(0).permutation(total)
    k = min(total, 64)
    while True:
        sample = np.sort(kb.upd[perm[:k]])
        t0 = time.perf_counter()
        idx.rebuild_update(sample)
        masks = idx.masks(plan)
        cnt, _ = idx.match(plan, kb.edge_ok, masks=masks, threads=threads)
        t = time.perf_counter() - t0
        if log:
            log(f"cpu [S2] sample {k}/{total} update keys: {t:.2f}s, {int(cnt)} matches")
        if t >= target_s or k == total:
            return t_idx + t * total / k, k / total, int(cnt)
        k = min(total, k * 4)


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by a copy-kernel
    calibration), or None when no summary covers this workload."""
    import glob
    for path in glob.glob(os.path.join(ROOT, "profiles", "*.traffic.txt")):
        kv = {}
        for ln in open(path):
            if ":" in ln and not ln.startswith("#"):
                k, v = ln.split(":", 1)
                kv[k.strip()] = v.strip()
        if kv.get("kernel") == kernel and kv.get("workload") == workload:
            return int(kv["traffic_bytes_per_launch"])
    return None


# Registered queries of the multi-query comparison: (kind, vertices, seed) of
# seeded random walks on the same graph, uneven in size and shape.
MQ_MIX = (("dense", 5, 101), ("tree", 5, 101), ("sparse", 5, 103), ("dense", 4, 101),
          ("tree", 4, 100), ("sparse", 4, 104), ("dense", 6, 103), ("tree", 6, 104))


TUNE_GRABS = (1, 2, 4, 8, 16)
TUNE_BPC = (0, 2)


def tuned_batch(eng, qs, log):
    """Stage 2 of the staged batch of each query in `qs` with the final-DFS
    launch shape autotuned on it (cgbdsm_amd.tune); returns the chosen shape,
    the DFS time per candidate and the batch's counts."""
    from cgbdsm_amd.tune import autotune

    live = []
    for q in qs:
        q.transfer_update_all()
        if q.transfer_update_index():
            q.build_local_index()
            live.append(q)
    res = {"rows_per_grab": None, "blocks_per_cu": None, "dfs_ms": {}, "counts": [0] * len(qs)}
    if live:
        best, table, got = autotune(eng, live, grabs=TUNE_GRABS, blocks_per_cu=TUNE_BPC, log=log)
        res.update(rows_per_grab=best[0], blocks_per_cu=best[1],
                   dfs_ms={f"{g}/{b}": round(ms, 3) for (g, b), ms in table.items()})
        for q, (c, _) in zip(live, got):
            res["counts"][qs.index(q)] = c
        for q in live:
            q.batch_update_global_index()
    for q in qs:
        q.sync()
    return res


def multi_query(args, threads, log):
    """Several registered queries of uneven size over the same update stream,
    each with its own device query (index, stream).  Per batch:
      plain           the queries' Stage 2 runs one after another;
      served_threads  the queries' Stage 2 runs concurrently, one host thread
                      per query (separate streams, one DFS launch each);
      served          index maintenance and masks run concurrently per query,
                      then ONE final-DFS launch counts every query from a
                      shared work queue (cgbdsm_matching_batch), then the
                      commits.
    All modes consume identical batches from separate device queries, so
    every per-query count must agree."""
    from concurrent.futures import ThreadPoolExecutor

    from cgbdsm_amd.engine import Engine
    from cgbdsm_amd.host import HostSession, synth_spec
    from cgbdsm_amd.tune import autotune

    n = min(args.multi_queries, len(MQ_MIX))
    mix, sessions = [], []
    for kind, size, seed in MQ_MIX[:n]:
        spec = synth_spec(args.scale, args.avg_degree, args.labels, zipf_s=args.zipf, seed=1,
                          initial_frac=args.initial_frac, stream_edges=args.stream_edges, query_size=size,
                          query_kind=kind, query_seed=seed)
        sessions.append(HostSession(spec=spec, batch_edges=args.batch_edges, threads=threads))
        mix.append(f"{size}-{kind}-s{seed}")
    nb = min(s.num_batches for s in sessions)
    warm = max(0, min(args.warmup, nb - 1))
    steps = max(1, min(args.steps, nb - warm))
    eng = Engine(0)
    modes = ("plain", "served_threads", "served")
    sets = {m: [] for m in modes}
    for s in sessions:
        keys, dv, card, avg = s.initial_keys()
        plan = s.make_plan(card, avg)
        batches = [s.next_batch() for _ in range(warm + steps)]
        for name in sets:
            q = eng.query(s.query_tables)
            q.set_layout(args.layout)
            q.index_init(keys, dv)
            q.set_plan(plan)
            q.set_deadline(args.deadline_ms)
            for kb in batches:
                q.enqueue(kb)
            sets[name].append(q)
    log(f"multi-query: {n} queries {mix}, {warm} warmup + {steps} timed batches")
    pool = ThreadPoolExecutor(max_workers=n)
    tuned = None
    phases = []

    def one(q):
        q.stage_next()
        return q.run_batch()[0]

    def prepare(q):
        q.stage_next()
        q.transfer_update_all()
        if not q.transfer_update_index():
            return False
        q.build_local_index()
        q.build_frontier()
        return True

    res = {}
    for name, qs in sets.items():
        lat, counts = [], []
        for b in range(warm + steps):
            t0 = time.perf_counter()
            if name == "plain":
                out = [one(q) for q in qs]
            elif name == "served_threads":
                out = list(pool.map(one, qs))
            else:
                live = list(pool.map(prepare, qs))
                t1 = time.perf_counter()
                out = [0] * n
                ks = [k for k in range(n) if live[k]]
                if ks and b == 0 and warm > 0 and not args.no_autotune:
                    tuning = autotune(eng, [qs[k] for k in ks], grabs=TUNE_GRABS, blocks_per_cu=TUNE_BPC, log=log)
                    tuned = {"rows_per_grab": tuning[0][0], "blocks_per_cu": tuning[0][1]}
                    t1 = time.perf_counter()
                if ks:
                    lead = qs[ks[0]]
                    if b >= warm:
                        lead.profile(True)
                        lead.kernel_times(reset=True)
                    for k, (c, _) in zip(ks, eng.matching_batch([qs[k] for k in ks])):
                        out[k] = c
                    t2 = time.perf_counter()
                    list(pool.map(lambda k: qs[k].batch_update_global_index(), ks))
                for q in qs:
                    q.sync()
                if b >= warm and ks:
                    kt = lead.kernel_times()
                    lead.profile(False)
                    phases.append((t1 - t0, t2 - t1, time.perf_counter() - t2, kt["dfs"]["ms"],
                                   lead.stats()["last_dfs_busy"]))
            dt = 1e3 * (time.perf_counter() - t0)
            if b >= warm:
                lat.append(dt)
            counts.append(out)
        res[name] = (lat, counts)
        log(f"multi-query {name}: {statistics.mean(lat):.1f} ms per batch")
    pool.shutdown()
    pairs = args.batch_edges
    out = {"queries": n, "mix": mix, "timed_batches": steps,
           "counts_equal": all(res[m][1] == res["plain"][1] for m in modes),
           "matches": int(sum(sum(c) for c in res["plain"][1][warm:]))}
    for name, (lat, _) in res.items():
        out[name] = {"update_edges_per_s": round(pairs * len(lat) / (1e-3 * sum(lat)), 3),
                     "p50_batch_ms": round(percentile(lat, 50), 4), "p99_batch_ms": round(percentile(lat, 99), 4)}
    out["served_vs_plain"] = round(out["served"]["update_edges_per_s"] / out["plain"]["update_edges_per_s"], 4)
    out["served_threads_vs_plain"] = round(out["served_threads"]["update_edges_per_s"] /
                                           out["plain"]["update_edges_per_s"], 4)
    out["served_dfs_tuning"] = tuned
    if phases:
        # wall ms per batch: index maintenance + masks of all queries
        # (concurrent), batched matching (seeds + one DFS launch), commits
        out["served_phase_ms"] = {k: round(1e3 * statistics.mean(p[i] for p in phases), 3)
                                  for i, k in enumerate(("prepare", "matching", "commit"))}
        out["served_phase_ms"]["dfs_kernel"] = round(statistics.mean(p[3] for p in phases), 3)
        out["served_dfs_busy"] = round(statistics.mean(p[4] for p in phases), 4)
        log(f"multi-query served phases: {out['served_phase_ms']}")
    for qs in sets.values():
        for q in qs:
            q.close()
    eng.close()
    return out


def serve_bench(args, threads, log):
    """Serving loop (cgbdsm_amd.serve) on one update stream: batch sizes drawn
    lognormal (median --serve-median edges, sigma 1, capped at 2 M), the
    8-query mix, half registered at the start and the rest arriving during
    the run.  Open loop: batches arrive as a Poisson process at --serve-rate
    batches/s, registrations at a third of that rate; closed loop: every
    batch is waiting at the start (capacity).  Each run in both modes: served
    (concurrent index maintenance + one batched DFS launch) and plain (the
    live queries' Stage 2 one after another); every count must agree."""
    import random

    import numpy as np

    from cgbdsm_amd.engine import Engine
    from cgbdsm_amd.host import HostSession, synth_spec
    from cgbdsm_amd.serve import Registration, Server

    rng = random.Random(7)
    n = min(args.multi_queries or 8, len(MQ_MIX))
    sessions = []
    for kind, size, seed in MQ_MIX[:n]:
        spec = synth_spec(args.scale, args.avg_degree, args.labels, zipf_s=args.zipf, seed=1,
                          initial_frac=args.initial_frac, stream_edges=args.stream_edges, query_size=size,
                          query_kind=kind, query_seed=seed)
        sessions.append((f"{size}-{kind}-s{seed}", HostSession(spec=spec, threads=threads)))
    left = sessions[0][1].stream_edges_left()
    sizes = []
    while left > 0 and len(sizes) < args.serve_batches:
        k = min(left, max(1, min(2_000_000, int(round(args.serve_median * float(np.exp(rng.gauss(0.0, 1.0))))))))
        sizes.append(k)
        left -= k
    regs = []
    for name, h in sessions:
        keys, dv, card, avg = h.initial_keys()
        plan = h.make_plan(card, avg)
        batches = [h.next_batch(k) for k in sizes]
        regs.append(Registration(name, h.query_tables, plan, keys, dv, batches))
    pairs = [int(b.pairs) for b in regs[0].batches]
    log(f"serve: {len(sizes)} batches (median {int(np.median(pairs))} edges, max {max(pairs)}), {n} queries")
    t, arr = 0.0, []
    for p in pairs:
        arr.append((t, p))
        t += rng.expovariate(args.serve_rate)
    t, reg_t = 0.0, []
    for i in range(n):
        if i < n // 2:
            reg_t.append(0.0)
        else:
            t += rng.expovariate(args.serve_rate / 3.0)
            reg_t.append(t)
    out = {"metric": "serving loop: sustained update edges/s and p50/p99 per-batch latency, served vs plain "
                     "(exact counts)",
           "batches": len(sizes), "batch_edges": pairs, "queries": [r.name for r in regs],
           "registration_s": [round(x, 4) for x in reg_t], "rate_batches_per_s": args.serve_rate}
    counts, tuning = {}, None
    for loop in ("open", "closed"):
        for mode in ("plain", "served"):
            mix = [Registration(r.name, r.tables, r.plan, r.keys, r.dv, r.batches, arrive_s=(a if loop == "open"
                                                                                              else 0.0))
                   for r, a in zip(regs, reg_t)]
            arrivals = arr if loop == "open" else [(0.0, p) for p in pairs]
            eng = Engine(0)
            # the launch shape is tuned once (first served batch of the open
            # loop) and kept, as a deployment would
            srv = Server(eng, mode=mode, threads=min(16, threads), deadline_ms=args.deadline_ms, tuning=tuning)
            res = srv.run(mix, arrivals)
            if mode == "served":
                tuning = res["dfs_tuning"]
            srv.close()
            eng.close()
            key = f"{loop}_{mode}"
            counts[key] = res["counts"]
            out[key] = {"update_edges_per_s": round(res["update_edges_per_s"], 3),
                        "p50_batch_ms": round(res["p50_ms"], 3), "p99_batch_ms": round(res["p99_ms"], 3),
                        "joined_batch": {k: v.get("joined_batch") for k, v in res["admissions"].items()}}
            log(f"serve {key}: {res['update_edges_per_s']:.0f} update edges/s, p50 {res['p50_ms']:.1f} ms, "
                f"p99 {res['p99_ms']:.1f} ms")
    # a query's count for a batch is the same in every run that served it
    agree = True
    for name in out["queries"]:
        merged = {}
        for c in counts.values():
            for b, v in c[name].items():
                if merged.setdefault(b, v) != v:
                    agree = False
    out["counts_agree"] = agree
    out["open_served_vs_plain"] = round(out["open_served"]["update_edges_per_s"] /
                                        out["open_plain"]["update_edges_per_s"], 4)
    out["closed_served_vs_plain"] = round(out["closed_served"]["update_edges_per_s"] /
                                          out["closed_plain"]["update_edges_per_s"], 4)
    return out


def index_stress(args, threads, log):
    """Index maintenance alone (apply, update index, commit; no matching),
    GAMMA lists vs the GPMA array on the same stream; the final main index
    of both layouts must be identical."""
    from cgbdsm_amd.engine import Engine
    from cgbdsm_amd.host import HostSession, synth_spec

    spec = synth_spec(args.scale, args.avg_degree, args.labels, zipf_s=args.zipf, seed=1,
                      initial_frac=args.initial_frac, stream_edges=args.stream_edges,
                      query_size=args.query_size, query_kind=args.query_kind, query_seed=args.query_seed)
    t0 = time.perf_counter()
    host = HostSession(spec=spec, batch_edges=args.batch_edges, threads=threads)
    keys, dv, card, avg = host.initial_keys()
    plan = host.make_plan(card, avg)
    batches = [host.next_batch() for _ in range(host.num_batches)]
    log(f"index stress: {keys.size} initial entries, {len(batches)} batches, Stage 1 done in "
        f"{time.perf_counter() - t0:.1f}s")
    warm = max(0, min(args.warmup, len(batches) - 1))
    steps = max(1, min(args.steps, len(batches) - warm))
    eng = Engine(0)
    out, finals = {}, {}
    for layout in ("gamma", "gpma"):
        q = eng.query(host.query_tables)
        q.set_layout(layout)
        q.index_init(keys, dv)
        q.set_plan(plan)
        for kb in batches[: warm + steps]:
            q.enqueue(kb)
        lat = []
        for i in range(warm + steps):
            if i == warm:
                q.profile(True)
                q.kernel_times(reset=True)
            q.sync()
            tb = time.perf_counter()
            q.stage_next()
            q.transfer_update_all()
            q.transfer_update_index()
            q.batch_update_global_index()
            q.sync()
            if i >= warm:
                lat.append(1e3 * (time.perf_counter() - tb))
        kt = q.kernel_times()
        st = q.stats()
        finals[layout] = q.export(0)
        edges = sum(kb.pairs for kb in batches[warm: warm + steps])
        out[layout] = {"update_edges_per_s": round(edges / (1e-3 * sum(lat)), 3),
                       "ms_per_batch": round(statistics.mean(lat), 3),
                       "p99_batch_ms": round(percentile(lat, 99), 3),
                       "main_entries": int(finals[layout].size),
                       "kernels_ms": {k: round(v["ms"], 3) for k, v in kt.items() if v["ms"] > 0},
                       "kernels_gbs": {k: round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 3) for k, v in kt.items()
                                       if v["ms"] > 0 and v["bytes"]}}
        if layout == "gpma":
            out[layout]["rebuilds"] = st["gpma_rebuilds"]
            out[layout]["windows"] = st["gpma_windows"]
        log(f"index stress {layout}: {statistics.mean(lat):.1f} ms per batch")
        q.close()
    eng.close()
    import numpy as np
    return {"metric": "index maintenance update edges/s (apply + update index + commit), GAMMA vs GPMA",
            "unit": "update edges/s", "steps": steps, "warmup": warm,
            "config": {"workload": f"rmat-s{args.scale}-deg{args.avg_degree:g}-{args.labels}labels-zipf{args.zipf:g}, "
                                   f"{args.initial_frac:g} initial, {args.stream_edges} streamed in "
                                   f"{args.batch_edges}-edge batches, {args.query_size}-vertex {args.query_kind} query"},
            "gamma": out["gamma"], "gpma": out["gpma"],
            "lists_equal": bool(np.array_equal(finals["gamma"], finals["gpma"]))}


def numa_nodes():
    try:
        return max(1, len([d for d in os.listdir("/sys/devices/system/node") if d.startswith("node") and
                           d[4:].isdigit()]))
    except OSError:
        return 1


def reference_threads():
    """The reference CPU path's default thread count: hardware threads / NUMA
    nodes (index/calig.cpp:1092-1095)."""
    total = os.cpu_count() or 1
    return max(1, total // numa_nodes()), total, numa_nodes()


def config4(args, log):
    """Batched multi-query headline: one update stream, many registered
    queries of uneven size, Stage 1 and Stage 2 on the device, plain (one
    query after another) vs served (concurrent Stage 1 and index work, one
    batched final-DFS launch).  A query whose exact count cannot be taken
    within --probe-ms on the first batch (or does not fit 64 bits) is
    excluded by name before timing."""
    from cgbdsm_amd.engine import Engine
    from cgbdsm_amd.host import Workload, synth_spec
    from cgbdsm_amd.multi import QuerySet, sample_mix
    from cgbdsm_amd.tune import autotune

    t_setup = time.perf_counter()
    spec = synth_spec(args.scale, args.avg_degree, args.labels, zipf_s=args.zipf, seed=1,
                      initial_frac=args.initial_frac, stream_edges=args.stream_edges)
    w = Workload(spec)
    specs = sample_mix(w, args.queries, args.mix_seed)
    log(f"workload: {w.n} vertices, {w.edges.shape[0]} initial edges, {w.stream.size // 2} stream edges, "
        f"{len(specs)} queries, generated in {time.perf_counter() - t_setup:.1f}s")
    budget = int(args.budget_gb * 2**30)
    eng = Engine(0, budget_bytes=budget)
    g = eng.graph(w.labels, w.edges)
    g.set_stream(w.stream)
    # Plain copies of every query first (at most half the budget, so the
    # served copies of the kept ones fit beside them with room to grow).
    errors = {}
    t_att = time.perf_counter()
    plain = QuerySet(eng, g, specs, layout=args.layout, deadline_ms=args.probe_ms, errors=errors,
                     stop_bytes=budget // 2)
    for m in plain.members:
        m.q.set_dfs_search(args.wave_search > 0, args.wave_search or 256)
    attach_s = time.perf_counter() - t_att
    log(f"{len(plain.members)} device queries attached (CaLiG static filter + initial index on the GPU) in "
        f"{attach_s:.1f}s, {eng.memory()[0] / 2**30:.1f} GiB held")
    bi = 2 * args.batch_edges
    nb = (w.stream.size + bi - 1) // bi
    warm = max(2, min(args.warmup, nb - 2))
    steps = max(1, min(args.steps, nb - warm))
    # Batch 0: which counts can be taken (plain, each query with the probe deadline).
    g.apply_range(0, bi)
    t0 = time.perf_counter()
    c_plain = plain.run_plain(errors)
    probe_s = time.perf_counter() - t0
    plain.drop(set(errors))
    kept_specs = [m.spec for m in plain.members]
    # Served copies of the kept queries, attached to the graph after batch 0:
    # the lit set is the greatest fixpoint on the current graph, so they
    # start from the index the plain copies hold after committing batch 0.
    t_att = time.perf_counter()
    served = QuerySet(eng, g, kept_specs, layout=args.layout, deadline_ms=args.deadline_ms, errors=errors,
                      stop_bytes=int(budget * 0.85))
    for m in served.members:
        m.q.set_dfs_search(args.wave_search > 0, args.wave_search or 256)
    plain.drop(set(errors))
    attach_s += time.perf_counter() - t_att
    excluded = {k: v.split(":")[0].strip() + ": " + v.split(";")[0].split(":", 1)[-1].strip()[:80]
                for k, v in errors.items()}
    for name in errors:
        log(f"excluded {name}: {errors[name][:160]}")
    log(f"served copies of {len(served.members)} queries attached; {eng.memory()[0] / 2**30:.1f} GiB held")
    tuned = {}

    def tune(qs):
        if tuned or args.no_autotune:
            return
        best, table, _ = autotune(eng, qs, grabs=TUNE_GRABS, blocks_per_cu=TUNE_BPC, log=log)
        tuned.update(rows_per_grab=best[0], blocks_per_cu=best[1],
                     dfs_ms={f"{a}/{b}": round(ms, 3) for (a, b), ms in table.items()})

    for m in plain.members:
        m.q.set_deadline(args.deadline_ms)
    counts_equal = []
    for b in range(1, warm):
        g.apply_range(b * bi, bi)
        counts_equal.append(plain.run_plain() == served.run_served(tune=tune if b == 1 else None)[0])
    log(f"{len(plain.members)} queries kept, {len(errors)} excluded; batch 0 plain {probe_s:.2f}s")
    lat_plain, lat_served, lat_apply, dfs_ms, dfs_bytes, s1_ms, matches, edges = [], [], [], [], [], [], 0, 0
    phases = []
    for b in range(warm, warm + steps):
        ta = time.perf_counter()
        g.apply_range(b * bi, bi)
        t_apply = time.perf_counter() - ta
        edges += min(bi, w.stream.size - b * bi) // 2
        tp = time.perf_counter()
        cp = plain.run_plain()
        t_plain = time.perf_counter() - tp
        s1_ms.append(sum(m.q.calig_stats()["ms"] for m in plain.members))
        ts = time.perf_counter()
        cs, dfs = served.run_served(profile_lead=True)
        t_served = time.perf_counter() - ts
        phases.append(served.last_phases)
        counts_equal.append(cp == cs)
        matches += sum(cs)
        lat_apply.append(1e3 * t_apply)
        lat_plain.append(1e3 * (t_apply + t_plain))
        lat_served.append(1e3 * (t_apply + t_served))
        if dfs and dfs["launches"]:
            dfs_ms.append(dfs["ms"] / dfs["launches"])
            dfs_bytes.append(dfs["bytes"] / dfs["launches"])
        log(f"batch {b}: apply {1e3 * t_apply:.1f} ms, plain {lat_plain[-1]:.1f} ms, served {lat_served[-1]:.1f} ms"
            f", {sum(cs)} matches, counts equal {cp == cs}")
    value = edges / (1e-3 * sum(lat_served))
    plain_rate = edges / (1e-3 * sum(lat_plain))
    launch_ms = statistics.mean(dfs_ms) if dfs_ms else 0.0
    bpl = statistics.mean(dfs_bytes) if dfs_bytes else 0.0
    achieved = bpl / (launch_ms * 1e-3) / 1e9 if launch_ms else 0.0
    workload = (f"config 4: rmat-s{args.scale}-deg{args.avg_degree:g}-{args.labels}labels-zipf{args.zipf:g}, "
                f"90/10 split seed 1, {args.batch_edges}-edge batches, {args.queries} random-walk queries of 4-12 "
                f"vertices (1/3 each tree/sparse/dense, mix seed {args.mix_seed}), device Stage 1 + Stage 2")
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "update edges/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(statistics.mean(lat_served), 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": workload, "batch_edges": args.batch_edges, "queries": args.queries,
                   "queries_kept": len(plain.members), "layout": args.layout, "parallelism": "single GPU"},
        "p50_batch_ms": round(percentile(lat_served, 50), 4), "p99_batch_ms": round(percentile(lat_served, 99), 4),
        "plain": {"update_edges_per_s": round(plain_rate, 3), "p50_batch_ms": round(percentile(lat_plain, 50), 4),
                  "p99_batch_ms": round(percentile(lat_plain, 99), 4)},
        "served_vs_plain": round(value / plain_rate, 4),
        "graph_apply_ms": round(statistics.mean(lat_apply), 4),
        # Served batch after the graph apply: Stage 1, index maintenance,
        # masks and frontiers of every query (concurrent), the batched launch,
        # the commits (concurrent); wall ms, mean over the timed batches.
        "served_phase_ms": {k: round(1e3 * statistics.mean(p[i] for p in phases), 3)
                            for i, k in enumerate(("prepare", "matching", "commit"))},
        "stage1_device_ms_per_batch_plain": round(statistics.mean(s1_ms), 3),
        "matches": int(matches), "matches_per_s": round(matches / (1e-3 * sum(lat_served)), 3),
        "counts_equal_plain_served": all(counts_equal),
        "excluded": excluded, "queries_kept": plain.names(), "attach_s": round(attach_s, 2),
        "dfs_tuning": tuned or None,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic("dfs_batched", workload),
                     "kernel": "KDfs (batched launch)", "launch_ms": round(launch_ms, 4),
                     "algorithmic_bytes_per_launch": int(bpl)},
    }
    kept = [m.spec for m in plain.members]
    plain.close()
    served.close()
    g.close()
    eng.close()
    if not args.no_cpu_baseline:
        cb = cpu_baseline_config4(args, kept, w, log)
        out["cpu_baseline"] = cb
        # The reference's plain GPU path runs Stage 1 (CaLiG) on the host for
        # each query: this run's plain batches with the device Stage-1 time
        # replaced by the host Stage-1 time measured on the CPU sample.
        host_s1 = cb["s1_ms_per_query"] * len(kept)
        ref_ms = statistics.mean(lat_plain) - statistics.mean(s1_ms) + host_s1
        out["plain_host_stage1_estimate"] = {
            "update_edges_per_s": round(args.batch_edges / (1e-3 * ref_ms), 3),
            "ms_per_batch": round(ref_ms, 3),
            "method": f"plain ms per batch - device Stage-1 ms + {len(kept)} x host Stage-1 ms per query "
                      f"({cb['s1_ms_per_query']} ms, mean of the CPU sample)"}
    w.close()
    return out


def cpu_baseline_config4(args, kept, w, log):
    """The CPU path (host CaLiG Stage 1 + the oracle's Stage 2, the reference's
    thread count) on a bounded sample: the first batch of the stream for the
    first kept query of each kind; the job's rate is the batch's edges over
    the sample's mean per-query time times the number of kept queries."""
    from cgbdsm_amd.host import HostSession, synth_spec

    threads, total, nodes = reference_threads()
    sample, per_q, s1_tot, s2_tot = [], [], 0.0, 0.0
    seen = set()
    for sp in kept:
        kind = sp.name.split("-")[1].rstrip("0123456789")
        if kind in seen:
            continue
        seen.add(kind)
        size = int(sp.name.split("-")[1][len(kind):])
        qseed = int(sp.name.split("-s")[-1])
        spec = synth_spec(args.scale, args.avg_degree, args.labels, zipf_s=args.zipf, seed=1,
                          initial_frac=args.initial_frac, stream_edges=args.stream_edges, query_size=size,
                          query_kind=kind, query_seed=qseed)
        t0 = time.perf_counter()
        host = HostSession(spec=spec, batch_edges=args.batch_edges, threads=threads)
        keys, dv, card, avg = host.initial_keys()
        plan = host.make_plan(card, avg)
        setup = time.perf_counter() - t0
        t1 = time.perf_counter()
        kb = host.next_batch()
        s1 = time.perf_counter() - t1
        log(f"cpu baseline {sp.name}: setup {setup:.1f}s, [S1] {1e3 * s1:.0f} ms")
        # (half the CPU budget: the sampled keys' costs are heavy-tailed, and
        # smaller samples gave estimates 3x apart between runs)
        s2, frac, _ = cpu_s2_sampled(keys, dv, plan, kb, threads, args.cpu_seconds / 2, log)
        sample.append(f"{sp.name} (setup {setup:.1f}s untimed, [S2] from {100 * frac:.3g} % of the update keys)")
        per_q.append(s1 + s2)
        s1_tot += s1
        s2_tot += s2
        log(f"cpu baseline {sp.name}: [S1] {1e3 * s1:.0f} ms, CPU [S2] {1e3 * s2:.0f} ms (estimated)")
        host.close()
        if s1_tot + s2_tot > args.cpu_seconds:
            break
    mean = statistics.mean(per_q)
    return {"value": round(args.batch_edges / (mean * len(kept)), 3), "unit": "update edges/s",
            "cores": threads, "nproc": total, "numa_nodes": nodes, "kind": "port",
            "sample": f"batch 0 ({args.batch_edges} edges) for {len(per_q)} of the {len(kept)} kept queries: "
                      + ", ".join(sample) + f"; [S2] = index maintenance + the masks and count of a uniform sample "
                      f"of the update keys, scaled by the inverse fraction; job rate = batch edges / "
                      f"(mean per-query time x {len(kept)})",
            "s1_ms_per_query": round(1e3 * s1_tot / len(per_q), 3),
            "s2_ms_per_query": round(1e3 * s2_tot / len(per_q), 3)}


def serve_device_bench(args, log):
    """Serving loop with Stage 1 on the device (cgbdsm_amd.serve.DeviceServer)
    on config 5: R-MAT s23, degree 24, 32 labels Zipf(1.1), 90 % of the
    edges resident; update batches arrive as a Poisson process with
    lognormal sizes (median --serve-median edges, sigma 1, capped at 2 M);
    queries of 4-12 vertices (mixed kinds) register at a tenth of the batch
    rate on top of --serve-initial registered at the start, at most 128
    live, admitted against the device memory budget.  Four runs over the
    same arrivals: open loop and closed loop (every batch waiting at the
    start), each served and plain; every count a query gets must agree
    across the runs.  A probe over the first three batches excludes, by
    name, the queries whose counts cannot be taken within the deadline (or
    exceed 64 bits); a later miss evicts the query, by name, in every run."""
    import random

    import numpy as np

    from cgbdsm_amd.engine import Engine
    from cgbdsm_amd.host import Workload, synth_spec
    from cgbdsm_amd.multi import sample_mix
    from cgbdsm_amd.serve import DeviceServer, QueryRequest

    rng = random.Random(7)
    t0 = time.perf_counter()
    spec = synth_spec(args.scale, args.avg_degree, args.labels, zipf_s=args.zipf, seed=1,
                      initial_frac=args.initial_frac, stream_edges=args.stream_edges)
    w = Workload(spec)
    left = w.stream.size // 2
    sizes = []
    while left > 0 and len(sizes) < args.serve_batches:
        k = min(left, max(1, min(2_000_000, int(round(args.serve_median * float(np.exp(rng.gauss(0.0, 1.0))))))))
        sizes.append(k)
        left -= k
    t, arr = 0.0, []
    for p in sizes:
        arr.append((t, p))
        t += rng.expovariate(args.serve_rate)
    horizon = t
    t, reg_t = 0.0, [0.0] * args.serve_initial
    while True:
        t += rng.expovariate(args.serve_rate / 10.0)
        if t >= horizon:
            break
        reg_t.append(t)
    specs = sample_mix(w, len(reg_t), args.mix_seed + 100)
    log(f"serve: graph of {w.n} vertices, {w.edges.shape[0]} resident edges; {len(sizes)} batches "
        f"(median {int(np.median(sizes))} edges, max {max(sizes)}), {len(specs)} registrations "
        f"({args.serve_initial} at start), setup {time.perf_counter() - t0:.1f}s")
    out = {"metric": "serving loop with device Stage 1: sustained update edges/s and p50/p99 per-batch latency, "
                     "served vs plain (exact counts)",
           "config": {"workload": f"config 5: rmat-s{args.scale}-deg{args.avg_degree:g}-{args.labels}labels-"
                                  f"zipf{args.zipf:g}, {args.initial_frac:g} resident, Poisson batches at "
                                  f"{args.serve_rate:g}/s, lognormal sizes (median {args.serve_median}, sigma 1, "
                                  f"cap 2M), registrations at a tenth of the batch rate, <= 128 live"},
           "batches": len(sizes), "batch_edges": sizes, "registrations": [sp.name for sp in specs],
           "registration_s": [round(x, 4) for x in reg_t]}
    # Probe: the first three batches, plain, every registration from the
    # start; a query whose count misses the deadline (or exceeds 64 bits)
    # there is excluded by name from the four timed runs, so their latency
    # measures serving, not evictions.  Later misses still evict.
    eng = Engine(0, budget_bytes=int(args.serve_budget_gb * (1 << 30)))
    g = eng.graph(w.labels, w.edges)
    g.set_stream(w.stream)
    srv = DeviceServer(eng, g, mode="plain", deadline_ms=args.serve_deadline_ms, threads=16, max_live=128,
                       log=lambda m: log(f"serve probe {m}"))
    probe = srv.run([QueryRequest(sp.name, sp.labels, sp.edges, arrive_s=0.0) for sp in specs],
                    [(0.0, p) for p in sizes[:3]])
    srv.close()
    g.close()
    eng.close()
    excluded = {n: v["reason"].split(";")[0][:120] for n, v in probe["evicted"].items()}
    keep = [i for i, sp in enumerate(specs) if sp.name not in excluded]
    specs, reg_t = [specs[i] for i in keep], [reg_t[i] for i in keep]
    out["excluded"] = excluded
    out["registrations_kept"] = [sp.name for sp in specs]
    log(f"serve probe: {len(excluded)} of {len(keep) + len(excluded)} registrations excluded")
    counts, tuning = {}, None
    for loop in ("open", "closed"):
        for mode in ("served", "plain"):
            eng = Engine(0, budget_bytes=int(args.serve_budget_gb * (1 << 30)))
            g = eng.graph(w.labels, w.edges)
            g.set_stream(w.stream)
            reqs = [QueryRequest(sp.name, sp.labels, sp.edges, arrive_s=(a if loop == "open" else min(a, 0.0)))
                    for sp, a in zip(specs, reg_t)]
            arrivals = arr if loop == "open" else [(0.0, p) for p in sizes]
            srv = DeviceServer(eng, g, mode=mode, deadline_ms=args.serve_deadline_ms, tuning=tuning,
                               threads=16, max_live=128, log=lambda m, k=f"{loop}_{mode}": log(f"serve {k} {m}"))
            res = srv.run(reqs, arrivals)
            if mode == "served":
                tuning = res["dfs_tuning"]
            srv.close()
            g.close()
            eng.close()
            key = f"{loop}_{mode}"
            counts[key] = res["counts"]
            out[key] = {"update_edges_per_s": round(res["update_edges_per_s"], 3),
                        "p50_batch_ms": round(res["p50_ms"], 3), "p99_batch_ms": round(res["p99_ms"], 3),
                        "evicted": res["evicted"], "admitted": len(res["admissions"]),
                        "footprint_mb": round((res["footprint_bytes"] or 0) / 2**20, 1)}
            log(f"serve {key}: {res['update_edges_per_s']:.0f} update edges/s, p50 {res['p50_ms']:.1f} ms, "
                f"p99 {res['p99_ms']:.1f} ms, evicted {list(res['evicted'])}")
    agree = True
    for name in out["registrations"]:
        merged = {}
        for c in counts.values():
            for b, v in c.get(name, {}).items():
                if merged.setdefault(b, v) != v:
                    agree = False
    out["counts_agree"] = agree
    out["open_served_vs_plain"] = round(out["open_served"]["update_edges_per_s"] /
                                        out["open_plain"]["update_edges_per_s"], 4)
    out["closed_served_vs_plain"] = round(out["closed_served"]["update_edges_per_s"] /
                                          out["closed_plain"]["update_edges_per_s"], 4)
    w.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--single", action="store_true",
                    help="plain single-query Stage-2 line (host Stage 1; R-MAT s20 dense query) instead of the "
                         "batched multi-query headline")
    ap.add_argument("--scale", type=int, default=None)
    ap.add_argument("--avg-degree", type=float, default=None)
    ap.add_argument("--labels", type=int, default=16)
    ap.add_argument("--zipf", type=float, default=1.2)
    ap.add_argument("--batch-edges", type=int, default=None)
    ap.add_argument("--initial-frac", type=float, default=0.9)
    ap.add_argument("--stream-edges", type=int, default=0, help="edges in the update stream (0: all the rest)")
    ap.add_argument("--queries", type=int, default=64, help="registered queries of the headline mix")
    ap.add_argument("--budget-gb", type=float, default=240.0, help="device memory budget of the headline (GiB)")
    ap.add_argument("--mix-seed", type=int, default=1, help="seed of the headline query mix")
    ap.add_argument("--probe-ms", type=float, default=1000.0,
                    help="headline: a query whose exact count of the first batch takes longer is excluded")
    ap.add_argument("--query-size", type=int, default=6)
    ap.add_argument("--query-kind", default="dense", choices=["dense", "tree", "sparse"])
    ap.add_argument("--query-seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--multi-only", action="store_true",
                    help="run only the 8-query comparison with host Stage 1 (plain / served_threads / served)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default final-DFS launch shape instead of tuning it on the first warmup batch")
    ap.add_argument("--deadline-ms", type=float, default=60_000)
    ap.add_argument("--donate-min", type=int, default=4, help="survivors left before a DFS wavefront donates (0: off)")
    ap.add_argument("--bfs-row-limit", type=int, default=65536)
    ap.add_argument("--no-shortest-list", action="store_true", help="final DFS scans the first backward neighbour")
    ap.add_argument("--no-product", action="store_true", help="final DFS enumerates the product suffix")
    ap.add_argument("--wave-search", type=int, default=64,
                    help="final DFS searches shared lists with the whole wavefront: at most 64 entries held one "
                         "per lane, this many or more (a power of two >= 64) by pivots; 0: off")
    ap.add_argument("--cpu-split", type=int, default=0,
                    help="--single: host threads counting a share of each final-DFS frontier beside the GPU "
                         "(0: off; -1: hardware threads / NUMA nodes)")
    ap.add_argument("--split-ratio", type=float, default=0.0,
                    help="--cpu-split: initial host share (0: one row per host thread, then the measured balance)")
    ap.add_argument("--layout", default="gpma", choices=["gamma", "gpma"],
                    help="main-index layout (both give identical lists; GPMA maintains them faster)")
    ap.add_argument("--serve", action="store_true",
                    help="serving loop, open and closed loop, served vs plain (prints its own JSON line)")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="batch arrivals per second (open loop)")
    ap.add_argument("--serve-median", type=int, default=50_000, help="median batch size in edges")
    ap.add_argument("--serve-batches", type=int, default=24, help="batches drawn from the stream at most")
    ap.add_argument("--serve-device", action="store_true",
                    help="config 5 serving loop with device Stage 1 (R-MAT s23; prints its own JSON line)")
    ap.add_argument("--serve-initial", type=int, default=32, help="queries registered before the first batch")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0, help="device memory budget of the serving runs")
    ap.add_argument("--serve-deadline-ms", type=float, default=3000.0,
                    help="per-batch count deadline; a query that misses it is evicted")
    ap.add_argument("--index-stress", action="store_true",
                    help="index maintenance only, GAMMA vs GPMA (prints its own JSON line)")
    ap.add_argument("--multi-queries", type=int, default=8,
                    help="queries in the host-Stage-1 multi-query comparison of --single; 0 skips it")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: {max(args.gpus, world)} GPUs requested; this build measures one GPU only "
              "(the path would shard by frontier rows with one u64 all-reduce per batch; no multi-GPU run "
              "is implemented).", file=sys.stderr)
        sys.exit(2)

    from cgbdsm_amd.engine import Engine
    from cgbdsm_amd.host import HostSession, synth_spec

    threads = min(16, os.cpu_count() or 1)
    headline = not (args.single or args.index_stress or args.serve or args.multi_only or args.serve_device)
    if args.serve_device:
        args.scale = args.scale or 23
        args.avg_degree = args.avg_degree or 24
        if args.labels == 16:
            args.labels = 32
        if args.zipf == 1.2:
            args.zipf = 1.1
        print(json.dumps(serve_device_bench(args, log)), flush=True)
        return
    if args.scale is None:
        args.scale = 21 if headline else 20
    if args.avg_degree is None:
        args.avg_degree = 24 if headline else 16
    if args.batch_edges is None:
        args.batch_edges = 250_000 if headline else 100_000
    if args.index_stress:
        print(json.dumps(index_stress(args, threads, log)), flush=True)
        return
    if args.serve:
        print(json.dumps(serve_bench(args, threads, log)), flush=True)
        return
    if args.multi_only:
        out = multi_query(args, threads, log)
        out["metric"] = "update edges/s, served multi-query vs plain (exact counts)"
        print(json.dumps(out), flush=True)
        return
    if headline:
        print(json.dumps(config4(args, log)), flush=True)
        return
    t_setup = time.perf_counter()
    spec = synth_spec(args.scale, args.avg_degree, args.labels, zipf_s=args.zipf, seed=1,
                      initial_frac=args.initial_frac, stream_edges=args.stream_edges,
                      query_size=args.query_size, query_kind=args.query_kind, query_seed=args.query_seed)
    host = HostSession(spec=spec, batch_edges=args.batch_edges, threads=threads)
    keys, dv, card, avg = host.initial_keys()
    plan = host.make_plan(card, avg)
    qlab, qedges = host.query_info()
    log(f"graph + CaLiG ready in {time.perf_counter() - t_setup:.1f}s: {keys.size} initial index entries, "
        f"dv {dv}, query labels {qlab.tolist()} edges {qedges.reshape(-1, 2).tolist()}, {host.num_batches} batches")

    # Stage 1 for every batch (host CPU), before timing.
    batches, s1_ms = [], []
    for _ in range(host.num_batches):
        kb = host.next_batch()
        batches.append(kb)
        s1_ms.append(sum(kb.s1_ms.values()))
    warm = max(0, min(args.warmup, len(batches) - 1))
    steps = max(1, min(args.steps, len(batches) - warm))
    if steps != args.steps:
        log(f"stream has {len(batches)} batches: timing {steps} steps after {warm} warmup")

    eng = Engine(0)
    q = eng.query(host.query_tables)
    q.set_layout(args.layout)
    q.index_init(keys, dv)
    q.set_plan(plan)
    q.set_deadline(args.deadline_ms)
    q.set_match_params(bfs_row_limit=args.bfs_row_limit, donate_min=args.donate_min)
    q.set_dfs_options(shortest_list=not args.no_shortest_list, product_suffix=not args.no_product)
    q.set_dfs_search(args.wave_search > 0, args.wave_search or 256)
    if args.cpu_split:
        q.set_cpu_split(threads=args.cpu_split, ratio=args.split_ratio, alpha=0.3, min_rows=0)
    split_log = []
    enq_ms = []
    for kb in batches[: warm + steps]:
        t_e = time.perf_counter()
        q.enqueue(kb)
        q.sync()
        enq_ms.append(1e3 * (time.perf_counter() - t_e))

    counts, lat_ms = [], []
    tuning = None
    for i in range(warm):
        q.stage_next()
        if i == 0 and not args.no_autotune:
            tuning = tuned_batch(eng, [q], log)
            counts.append(tuning.pop("counts")[0])
            continue
        c, _, _ = q.run_batch()
        counts.append(c)
    q.profile(True)
    q.kernel_times(reset=True)
    split_before = q.split_stats()["batches"] if args.cpu_split else 0
    q.sync()
    t0 = time.perf_counter()
    stage_ms = {"apply": 0.0, "update_index": 0.0, "masks": 0.0, "match": 0.0, "commit": 0.0}
    dfs = {"rows": 0, "scanned": 0, "probes": 0, "donated": 0}
    for i in range(steps):
        tb = time.perf_counter()
        q.stage_next()
        c, _, _ = q.run_batch()
        lat_ms.append(1e3 * (time.perf_counter() - tb))
        counts.append(c)
        st = q.stats()
        for k in stage_ms:
            stage_ms[k] += st["ms_" + k]
        dfs["rows"] += st["last_frontier_rows"]
        dfs["scanned"] += st["last_dfs_scanned"]
        dfs["probes"] += st["last_dfs_probes"]
        dfs["donated"] += st["last_dfs_donated"]
        log(f"batch {warm + i}: {c} matches in {lat_ms[-1]:.1f} ms, frontier {st['last_frontier_rows']} rows at "
            f"depth {st['last_frontier_depth']}, scanned {st['last_dfs_scanned']}, probes {st['last_dfs_probes']}, "
            f"donated {st['last_dfs_donated']}, busy {st['last_dfs_busy']:.3f}, steps {st['last_dfs_steps']}, "
            f"lanes/step {st['last_dfs_step_lanes'] / max(1, st['last_dfs_steps']):.1f}, "
            f"LDS cache {st['last_dfs_lds_cache']} u32/wave")
        sp = q.split_stats() if args.cpu_split else None
        if sp and sp["batches"] > (split_log[-1]["batches"] if split_log else split_before):
            split_log.append({"batch": warm + i, **{k: (round(v, 6) if isinstance(v, float) else v)
                                                     for k, v in sp.items()}})
            log(f"  split: share {sp['ratio']:.6f}, host {sp['last_cpu_rows']} rows / {sp['last_cpu_ms']:.1f} ms "
                f"({sp['last_cpu_matches']} matches) after {sp['last_wait_ms']:.1f} ms waiting for the mirror, "
                f"GPU {sp['last_gpu_rows']} rows / {sp['last_gpu_ms']:.1f} ms, "
                f"mirror {sp['mirror_entries']} entries, {sp['mirror_ms']:.1f} ms host upkeep")
    q.sync()
    elapsed = time.perf_counter() - t0
    kt = q.kernel_times()
    timed = batches[warm: warm + steps]
    edges = sum(kb.pairs for kb in timed)
    matches = sum(counts[warm:])
    value = edges / elapsed

    # Dominant kernel class by summed device time.
    dom = max(kt, key=lambda k: kt[k]["ms"])
    d = kt[dom]
    launch_ms = d["ms"] / max(d["launches"], 1)
    bytes_per_launch = d["bytes"] / max(d["launches"], 1)
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9 if launch_ms > 0 else 0.0
    split = f"{round(100 * args.initial_frac)}/{round(100 * (1 - args.initial_frac))}"
    if args.stream_edges:
        split += f" ({args.stream_edges} streamed)"
    workload = (f"rmat-s{args.scale}-deg{args.avg_degree:g}-{args.labels}labels-zipf{args.zipf:g} "
                f"{split} stream, {args.batch_edges}-edge batches, plain single query: "
                f"{args.query_size}-vertex {args.query_kind} random-walk query seed {args.query_seed}")
    roofline = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": measured_traffic(dom, workload), "kernel": dom,
                "launch_ms": round(launch_ms, 4), "algorithmic_bytes_per_launch": int(bytes_per_launch)}

    mq = multi_query(args, threads, log) if args.multi_queries > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        c_edges, c_secs, c_counts = cpu_baseline(keys, dv, plan, batches, args.cpu_seconds, threads)
        gpu_counts = counts[: len(c_counts)]
        cpu = {"value": round(c_edges / c_secs, 3) if c_secs > 0 else None, "unit": "update edges/s",
               "cores": threads, "kind": "port",
               "sample": f"first {len(c_counts)} of {len(batches)} batches of this stream (oracle Stage 2: index "
                         f"merges, masks, seeding + DFS with {threads} threads), {c_secs:.1f}s",
               "counts_match_gpu": gpu_counts == c_counts}
        if gpu_counts != c_counts:
            log(f"COUNT MISMATCH gpu {gpu_counts} cpu {c_counts}")

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "update edges/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "ms_per_step": round(1e3 * elapsed / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": workload,
                   "batch_edges": args.batch_edges, "query_edges": int(qedges.size // 2), "layout": args.layout,
                   "parallelism": "single GPU"},
        "p50_batch_ms": round(percentile(lat_ms, 50), 4), "p99_batch_ms": round(percentile(lat_ms, 99), 4),
        # Host keys -> HBM over PCIe (cgbdsm_batch_enqueue), outside the timed region.
        "enqueue_ms_per_batch": round(statistics.mean(enq_ms[warm:]), 4),
        "pcie_inclusive_update_edges_per_s": round(edges / (elapsed + 1e-3 * sum(enq_ms[warm:])), 3),
        "matches": int(matches), "matches_per_s": round(matches / elapsed, 3),
        "stage_ms_per_batch": {k: round(v / steps, 4) for k, v in stage_ms.items()},
        "s1_cpu_ms_per_batch": round(statistics.mean(s1_ms), 3),
        "kernels_ms": {k: round(v["ms"], 4) for k, v in kt.items()},
        "kernels_gbs": {k: round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 3) for k, v in kt.items() if v["ms"] > 0 and v["bytes"]},
        "dfs_totals": dfs,
        "roofline": roofline, "cpu_baseline": cpu, "multi_query": mq, "dfs_tuning": tuning,
        "cpu_split": {"threads": args.cpu_split, "final_share": round(q.split_stats()["ratio"], 9),
                      "split_batches": split_log} if args.cpu_split else None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
