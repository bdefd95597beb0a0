"""Autotuner of the final-DFS launch shape.

The knobs are the frontier rows a wavefront takes per cursor atomic (the
stealing chunk of the row phase) and the resident blocks per compute unit
(wavefronts in flight, each with its LDS stack of up to 12 levels).  The
tuner counts ONE batch once per candidate (cgbdsm_matching may count the
same batch again: the index is only read), times each final-DFS launch with
the HIP events the engine records on the query's stream, requires every
candidate to produce identical per-query, per-edge counts, and leaves the
fastest setting applied to the queries (event timing is switched off
again on the leading query).  No reference counterpart: the
reference's launch shape is fixed (utils/config.h GRID_DIM / BLOCK_DIM).
"""
import numpy as np

GRABS = (1, 2, 4, 8, 16)
BLOCKS_PER_CU = (0,)


def _count(engine, queries):
    if len(queries) == 1:
        c, per = queries[0].matching()
        return [(c, per)]
    return engine.matching_batch(queries)


def autotune(engine, queries, grabs=GRABS, blocks_per_cu=BLOCKS_PER_CU, repeats=1, log=None):
    """`queries`: device queries after build_local_index on the same batch
    (the first one leads a batched launch).  Returns (best, table, counts):
    best = (rows_per_grab, blocks_per_cu), table = {setting: DFS ms},
    counts = the batch's [(count, per_edge)] (identical for every setting)."""
    lead = queries[0]
    lead.profile(True)
    table, ref = {}, None
    for bpc in blocks_per_cu:
        for g in grabs:
            for q in queries:
                q.set_dfs_tuning(g, bpc)
            best_ms = None
            for _ in range(max(1, repeats)):
                lead.kernel_times(reset=True)
                got = _count(engine, queries)
                ms = lead.kernel_times()["dfs"]["ms"]
                best_ms = ms if best_ms is None else min(best_ms, ms)
                if ref is None:
                    ref = got
                elif any(a[0] != b[0] or not np.array_equal(a[1], b[1]) for a, b in zip(got, ref)):
                    raise RuntimeError(f"DFS tuning {(g, bpc)} changed the counts")
            table[(g, bpc)] = best_ms
            if log:
                log(f"autotune rows_per_grab={g} blocks_per_cu={bpc}: DFS {best_ms:.2f} ms")
    best = min(table, key=table.get)
    for q in queries:
        q.set_dfs_tuning(*best)
    lead.kernel_times(reset=True)
    lead.profile(False)
    return best, table, ref


# Second stage: stealing knobs of the final DFS, searched one at a time
# around the launch shape found first (coordinate descent: each candidate
# costs one counted batch).  range_min is the smallest candidate range a
# wavefront splits off for idle ones (the reference's fixed STEAL_CHUNK /
# TASK_SHARE_BOUND, kernels/enumeration_balance.cu:11-17); donate_min the
# survivors left at which it hands work to the queue; cache_cap the u32 per
# wavefront of the LDS stable-intersection cache (0: all the LDS left);
# wave_search the shortest list (a power of two >= 64) searched by the whole
# wavefront's pivots instead of one lane's binary search (0: never).
RANGE_MIN = (256, 1024, 4096)
DONATE_MIN = (1, 4, 16)
CACHE_CAP = (0, 1024, 256)
WAVE_SEARCH = (64, 256, 1024, 0)


def autotune_full(engine, queries, grabs=GRABS, blocks_per_cu=BLOCKS_PER_CU, range_min=RANGE_MIN,
                  donate_min=DONATE_MIN, cache_cap=CACHE_CAP, wave_search=WAVE_SEARCH, log=None):
    """autotune, then the stealing and search knobs one after another.
    Returns (knobs, table, counts): knobs = {rows_per_grab, blocks_per_cu,
    range_min, donate_min, cache_cap, wave_search}, table = {knob: {value: DFS
    ms}}; every candidate must give identical counts, and the fastest setting
    is left applied."""
    best, shape_table, ref = autotune(engine, queries, grabs=grabs, blocks_per_cu=blocks_per_cu, log=log)
    knobs = {"rows_per_grab": best[0], "blocks_per_cu": best[1], "range_min": 1024, "donate_min": 4,
             "cache_cap": 0, "wave_search": 64}
    table = {"shape": {f"{g}/{b}": ms for (g, b), ms in shape_table.items()}}
    lead = queries[0]

    def apply(k):
        for q in queries:
            q.set_dfs_steal(k["range_min"], k["cache_cap"])
            limit, cap, _ = getattr(q, "match_params", (65536, 0, 4))
            q.set_match_params(bfs_row_limit=limit, frontier_cap=cap, donate_min=k["donate_min"])
            q.set_dfs_search(k["wave_search"] > 0, k["wave_search"] or 64)

    lead.profile(True)
    for name, values in (("range_min", range_min), ("donate_min", donate_min), ("cache_cap", cache_cap),
                         ("wave_search", wave_search)):
        table[name] = {}
        for v in values:
            trial = dict(knobs, **{name: v})
            apply(trial)
            lead.kernel_times(reset=True)
            got = _count(engine, queries)
            ms = lead.kernel_times()["dfs"]["ms"]
            if any(a[0] != b[0] or not np.array_equal(a[1], b[1]) for a, b in zip(got, ref)):
                raise RuntimeError(f"DFS knob {name}={v} changed the counts")
            table[name][v] = ms
            if log:
                log(f"autotune {name}={v}: DFS {ms:.2f} ms")
        knobs[name] = min(table[name], key=table[name].get)
    apply(knobs)
    lead.kernel_times(reset=True)
    lead.profile(False)
    return knobs, table, ref
