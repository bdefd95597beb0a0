"""Serving loops (cgbdsm_amd.serve): queries registered while batches are
being served, admitted against the device memory budget, brought up to the
stream position on their own stream and counted in the batched launch.
Every count a query gets in the mix must equal its count when it runs alone
from the start of the stream, and the oracle's count of that batch.  The
device-Stage-1 server (DeviceServer) attaches late queries over the
resident graph as it stands; its counts must equal the oracle's too."""
import numpy as np
import pytest

from cgbdsm_amd.engine import Engine
from cgbdsm_amd.host import HostSession, Workload, synth_spec
from cgbdsm_amd.serve import DeviceServer, QueryRequest, Registration, Server
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu

MIX = [("dense", 5, 2), ("tree", 5, 3), ("sparse", 6, 4), ("dense", 4, 5)]


def _oracle_counts(plan, keys, dv, batches):
    idx = O.Index(plan.ndir, 5)
    idx.set_dv(dv)
    idx.init(keys)
    out = {}
    for b, kb in enumerate(batches):
        idx.set_dv(kb.dv)
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        idx.rebuild_update(kb.upd)
        c = 0
        if kb.upd.size:
            c, _ = idx.match(plan, kb.edge_ok, masks=idx.masks(plan), threads=4)
            idx.insert(kb.upd)
        out[b] = int(c)
    return out


ORACLE = {}


@pytest.fixture(scope="module")
def regs():
    out = []
    for i, (kind, size, qseed) in enumerate(MIX):
        spec = synth_spec(11, 8, 4, seed=2, query_size=size, query_kind=kind, query_seed=qseed)
        host = HostSession(spec=spec, batch_edges=120, threads=4)
        keys, dv, card, avg = host.initial_keys()
        plan = host.make_plan(card, avg)
        batches = [host.next_batch() for _ in range(host.num_batches)]
        r = Registration(f"q{i}-{size}{kind}", host.query_tables, plan, keys, dv, batches)
        ORACLE[r.name] = _oracle_counts(plan, keys, dv, batches)
        r.query = host.query_info()
        out.append(r)
    return out


def _alone(eng, r, nb):
    dq = eng.query(r.tables)
    dq.set_layout("gpma")
    dq.index_init(r.keys, r.dv)
    dq.set_plan(r.plan)
    counts, peak = {}, 0
    for b in range(nb):
        dq.stage(r.batches[b])
        c, _, _ = dq.run_batch()
        counts[b] = c
        peak = max(peak, eng.memory()[0])
    dq.close()
    return counts, peak


@pytest.fixture(scope="module")
def alone(regs):
    eng = Engine(0)
    nb = min(len(r.batches) for r in regs)
    res = {r.name: _alone(eng, r, nb) for r in regs}
    eng.close()
    for r in regs:  # alone on the device == the oracle
        for b in range(nb):
            assert res[r.name][0][b] == ORACLE[r.name][b], (r.name, b)
    return nb, res


def _with_times(regs, arrive, leave):
    return [Registration(r.name, r.tables, r.plan, r.keys, r.dv, r.batches, arrive_s=a, leave_after=l)
            for r, a, l in zip(regs, arrive, leave)]


@pytest.mark.parametrize("mode", ["served", "plain"])
def test_counts_in_the_mix_equal_alone(regs, alone, mode):
    print(f"case serve mix {mode}", flush=True)
    nb, ref = alone
    assert nb >= 6
    arrivals = [(0.03 * b, int(regs[0].batches[b].pairs)) for b in range(nb)]
    # two queries from the start (one leaves after batch 3), two arrive later
    mix = _with_times(regs, [0.0, 0.0, 0.07, 0.13], [None, 3, None, nb - 2])
    eng = Engine(0)
    srv = Server(eng, mode=mode, threads=8)
    res = srv.run(mix, arrivals)
    srv.close()
    eng.close()
    served = 0
    for r in mix:
        got = res["counts"][r.name]
        assert got, f"{r.name} was never served"
        for b, c in got.items():
            assert c == ref[r.name][0][b], (r.name, b, c, ref[r.name][0][b])
            served += 1
        if r.leave_after is not None:
            assert max(got) == r.leave_after
    assert res["admissions"][mix[2].name]["joined_batch"] > 0
    assert res["admissions"][mix[3].name]["joined_batch"] > 0
    assert len(res["latency_ms"]) == nb and served >= 2 * nb
    print(f"served {served} query-batches, p50 {res['p50_ms']:.2f} ms, p99 {res['p99_ms']:.2f} ms", flush=True)


def test_admission_waits_for_memory(regs, alone):
    """Under a budget that holds two of these queries, the third is admitted
    only after one leaves; its counts still equal the alone run."""
    print("case serve budget", flush=True)
    nb, ref = alone
    same = [regs[1]] * 3  # the same query three times: equal footprints
    peak = ref[regs[1].name][1]
    budget = int(2.5 * peak)
    mix = [Registration(f"t{i}", r.tables, r.plan, r.keys, r.dv, r.batches, arrive_s=0.0,
                        leave_after=(2 if i == 0 else None), footprint=peak) for i, r in enumerate(same)]
    arrivals = [(0.03 * b, int(same[0].batches[b].pairs)) for b in range(nb)]
    eng = Engine(0, budget_bytes=budget)
    srv = Server(eng, threads=8)
    res = srv.run(mix, arrivals)
    srv.close()
    eng.close()
    adm = res["admissions"]
    assert adm["t1"]["at_batch"] == 0
    assert adm["t2"]["at_batch"] >= 3 and adm["t2"]["joined_batch"] >= 3  # after t0 left (after batch 2)
    for r in mix:
        for b, c in res["counts"][r.name].items():
            assert c == ref[regs[1].name][0][b]
    assert np.all(np.diff(sorted(res["counts"]["t2"])) == 1)


@pytest.mark.parametrize("mode", ["served", "plain"])
def test_device_server_counts_equal_the_oracle(regs, alone, mode):
    """Stage 1 on the device: the graph and stream resident, two queries from
    the start (one leaves after batch 3), two attached over the graph at
    later batch boundaries (each joins at the batch after its admission);
    every count equals the oracle's count of that batch."""
    print(f"case device serve {mode}", flush=True)
    nb, _ = alone
    w = Workload(synth_spec(11, 8, 4, seed=2))
    eng = Engine(0)
    g = eng.graph(w.labels, w.edges)
    g.set_stream(w.stream)
    arrivals = [(0.03 * b, 120) for b in range(nb)]
    reqs = [QueryRequest(r.name, r.query[0], r.query[1].reshape(-1, 2), arrive_s=a, leave_after=l)
            for r, a, l in zip(regs, [0.0, 0.0, 0.05, 0.10], [None, 3, None, None])]
    srv = DeviceServer(eng, g, mode=mode, threads=8)
    res = srv.run(reqs, arrivals)
    srv.close()
    g.close()
    eng.close()
    w.close()
    served = 0
    for r in reqs:
        got = res["counts"][r.name]
        assert got, f"{r.name} was never served"
        for b, c in got.items():
            assert c == ORACLE[r.name][b], (r.name, b, c, ORACLE[r.name][b])
            served += 1
    assert res["admissions"][reqs[2].name]["joined_batch"] > 0
    assert res["admissions"][reqs[3].name]["joined_batch"] > 0
    assert max(res["counts"][reqs[1].name]) == 3
    assert not res["evicted"]
    assert served >= 2 * nb
