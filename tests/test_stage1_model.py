"""The device Stage-1 algorithm (csrc/hip/stage1.hip), restated as a plain
Python model and run on the CPU: static filter, then per batch the
dead-mask-driven activation closure and the deletion rounds over the
optimistically lit pairs.  Its lit bits must equal the oracle's from-scratch
greatest fixpoint after every batch (and the host producer's lit bits), on
streams where the fixpoint changes in every way (pairs lit by a new edge,
by chains and cycles of newly lit pairs, and pairs that fail only on
distinct representatives).  A model in which one rule is dropped must fail
on the same streams, so these streams exercise each rule.  On streams that
also delete edges, the model keeps the main index with the entries it
emits (new pairs, update keys, entries of deleted edges and of pairs that
die) and must equal the from-scratch index of the fixpoint every batch."""
import itertools

import numpy as np
import pytest

from cgbdsm_amd.host import HostSession, synth_spec
from tests import oracle_lib as O
from tests.parity import deletion_stream


class Model:
    def __init__(self, n, labels, adj, qlabels, qedges, drop=None):
        self.n, self.lab = n, labels
        self.adj = [set(a) for a in adj]
        self.qv = len(qlabels)
        self.ql = list(qlabels)
        self.qn = [set() for _ in range(self.qv)]
        for a, b in qedges:
            self.qn[a].add(b)
            self.qn[b].add(a)
        self.drop = drop
        self.lit = [0] * n
        self.dead = {}
        # constructCand: every compatible pair lit
        for v in range(n):
            for u in range(self.qv):
                if labels[v] == self.ql[u]:
                    self.lit[v] |= 1 << u
        pairs = [(v, u) for v in range(n) for u in range(self.qv) if (self.lit[v] >> u) & 1]
        self.rounds(pairs, opt=None)

    def supports(self, v, u):
        return {uj: [vj for vj in self.adj[v] if self.lab[vj] == self.ql[uj] and (self.lit[vj] >> uj) & 1]
                for uj in self.qn[u]}

    def check(self, v, u):
        s = self.supports(v, u)
        empty = [uj for uj, l in s.items() if not l]
        if empty:
            return False, sum(1 << uj for uj in empty)
        groups = {}
        for uj in self.qn[u]:
            groups.setdefault(self.ql[uj], []).append(uj)
        if self.drop != "injective":
            for g in groups.values():
                if len(g) > 1 and not any(len(set(c)) == len(g) for c in itertools.product(*[s[uj] for uj in g])):
                    return False, 0
        return True, 0

    def rounds(self, work, opt, opt_only=True):
        """Returns the pairs that died."""
        work = list(dict.fromkeys(work))
        gone = []
        while work:
            res = [(p, self.check(*p)) for p in work]  # all read the lit set of the round's start
            died = [(p, bits) for p, (ok, bits) in res if not ok]
            if not died:
                break
            for (v, u), bits in died:
                self.lit[v] &= ~(1 << u)
                if opt is not None:
                    opt.discard((v, u))
                self.dead[(v, u)] = bits
                gone.append((v, u))
            nxt = []
            for (v, u), _ in died:
                for uj in self.qn[u]:
                    for vj in self.adj[v]:
                        if self.lab[vj] == self.ql[uj] and (self.lit[vj] >> uj) & 1:
                            if opt is None or not opt_only or (vj, uj) in opt:
                                nxt.append((vj, uj))
            work = list(dict.fromkeys(nxt))
        return gone

    def edge(self, a, b):
        return a != b and 0 <= a < self.n and 0 <= b < self.n

    def batch(self, pairs):
        """One batch of (a, b) pairs: insertions, and deletions as pairs of
        negative ids (-a-1, -b-1); both judged against the graph before the
        batch.  Returns (inserted edges, deleted edges, Ins, Del, Upd) with
        the index entries as (ui, uj, v, w) tuples in original ids."""
        new, gone_e = [], []
        for a, b in pairs:
            if a < 0 and b < 0:
                x, y = -a - 1, -b - 1
                if self.edge(x, y) and y in self.adj[x] and (min(x, y), max(x, y)) not in gone_e:
                    gone_e.append((min(x, y), max(x, y)))
            elif self.edge(a, b) and b not in self.adj[a] and (min(a, b), max(a, b)) not in new:
                new.append((min(a, b), max(a, b)))
        lit0 = list(self.lit)
        for a, b in gone_e:
            self.adj[a].discard(b)
            self.adj[b].discard(a)
        for a, b in new:
            self.adj[a].add(b)
            self.adj[b].add(a)
        touched = []
        for a0, b0 in new:
            for a, b in ((a0, b0), (b0, a0)):
                for u in range(self.qv):
                    if self.lab[a] != self.ql[u] or (self.lit[a] >> u) & 1:
                        continue
                    m = sum(1 << uj for uj in self.qn[u] if self.ql[uj] == self.lab[b])
                    if m:
                        if self.drop != "edge_clear":
                            self.dead[(a, u)] = self.dead.get((a, u), 0) & ~m
                        touched.append((a, u))
        opt = set()
        while touched:
            act = []
            for v, u in dict.fromkeys(touched):
                if not (self.lit[v] >> u) & 1 and self.dead.get((v, u), 0) == 0:
                    self.lit[v] |= 1 << u
                    opt.add((v, u))
                    act.append((v, u))
            touched = []
            if self.drop == "push":
                break
            for v, u in act:
                for uj in self.qn[u]:
                    for vj in self.adj[v]:
                        if self.lab[vj] == self.ql[uj] and not (self.lit[vj] >> uj) & 1:
                            self.dead[(vj, uj)] = self.dead.get((vj, uj), 0) & ~(1 << u)
                            touched.append((vj, uj))
        # Deleted edges: the entries they carried, and the pairs from before
        # the batch that lost an edge join the check list.
        dele = []
        seeds = sorted(opt)
        for x, y in gone_e:
            for a, b in ((x, y), (y, x)):
                for ui in range(self.qv):
                    if not (lit0[a] >> ui) & 1:
                        continue
                    m = [uj for uj in self.qn[ui] if self.ql[uj] == self.lab[b]]
                    if m and self.drop != "del_seed":
                        seeds.append((a, ui))
                    if self.drop != "dedge_keys":
                        dele += [(ui, uj, a, b) for uj in m if (lit0[b] >> uj) & 1]
        gone = self.rounds(seeds, opt, opt_only=not gone_e)
        killed = {p for p in gone if (lit0[p[0]] >> p[1]) & 1}
        old = lambda v, u: bool((lit0[v] >> u) & 1) and ((self.lit[v] >> u) & 1 or (v, u) in killed)
        for v, ui in sorted(killed):
            for vj in self.adj[v]:
                for uj in self.qn[ui]:
                    if self.ql[uj] != self.lab[vj] or not old(vj, uj):
                        continue
                    if (vj, uj) in killed and vj < v or self.drop == "killed_keys":
                        continue
                    dele += [(ui, uj, v, vj), (uj, ui, vj, v)]
        ins = []
        for v, ui in sorted(opt):
            for vj in self.adj[v]:
                for uj in self.qn[ui]:
                    if self.ql[uj] == self.lab[vj] and (self.lit[vj] >> uj) & 1:
                        ins += [(ui, uj, v, vj), (uj, ui, vj, v)]
        upd = []
        for a, b in new:
            for v, w in ((a, b), (b, a)):
                for ui in range(self.qv):
                    if (self.lit[v] >> ui) & 1:
                        upd += [(ui, uj, v, w) for uj in self.qn[ui]
                                if self.ql[uj] == self.lab[w] and (self.lit[w] >> uj) & 1]
        return new, gone_e, ins, dele, upd

    def entries(self):
        """The main index from scratch: every edge between lit pairs."""
        return {(ui, uj, v, vj) for v in range(self.n) for ui in range(self.qv) if (self.lit[v] >> ui) & 1
                for vj in self.adj[v] for uj in self.qn[ui]
                if self.ql[uj] == self.lab[vj] and (self.lit[vj] >> uj) & 1}


def _csr(model):
    offs = np.zeros(model.n + 1, dtype=np.uint64)
    adj = []
    for v in range(model.n):
        nb = sorted(model.adj[v])
        adj.extend(nb)
        offs[v + 1] = offs[v] + len(nb)
    return offs, np.array(adj if adj else [0], dtype=np.uint32)


def _run(spec, be, drop=None, check_host=True):
    host = HostSession(spec=spec, batch_edges=be, threads=2)
    host.initial_keys()
    offs, adj, lab = host.graph()
    ql, qe = host.query_info()
    n = lab.size
    lists = [adj[int(offs[v]):int(offs[v + 1])].tolist() for v in range(n)]
    m = Model(n, lab.tolist(), lists, ql.tolist(), qe.reshape(-1, 2).tolist(), drop=drop)
    mism = 0
    lit = np.array(m.lit, dtype=np.uint32)
    if not np.array_equal(lit, O.li_fixpoint(offs, adj, lab, ql, qe)):
        mism += 1
    stream = host.update_stream()
    for b in range(host.num_batches):
        host.next_batch()
        m.batch(stream[b * be:(b + 1) * be].tolist())
        lit = np.array(m.lit, dtype=np.uint32)
        o2, a2 = _csr(m)
        fp = O.li_fixpoint(o2, a2, lab, ql, qe)
        if not np.array_equal(lit, fp):
            mism += 1
        if check_host and not np.array_equal(lit, host.lit(n)):
            mism += 1
    return mism


STREAMS = [
    (synth_spec(9, 6, 3, seed=1, query_size=5, query_kind="dense", query_seed=1), 120),
    (synth_spec(9, 6, 2, seed=2, query_size=6, query_kind="sparse", query_seed=2), 100),
    (synth_spec(9, 8, 2, seed=3, query_size=5, query_kind="tree", query_seed=3), 150),
    (synth_spec(8, 6, 1, seed=4, query_size=5, query_kind="dense", query_seed=4), 60),
]


@pytest.mark.parametrize("i", range(len(STREAMS)))
def test_model_reaches_the_fixpoint(i):
    spec, be = STREAMS[i]
    assert _run(spec, be) == 0


@pytest.mark.parametrize("drop", ["push", "edge_clear", "injective"])
def test_model_without_a_rule_fails(drop):
    """Each rule of the device algorithm is needed: without it, some stream
    ends at a different lit set than the fixpoint."""
    assert sum(_run(spec, be, drop=drop, check_host=False) for spec, be in STREAMS) > 0


def _run_deletions(spec, be, drop=None, seed=0):
    """Mixed insertion/deletion stream through the model: lit bits against
    the oracle's fixpoint, and the main index kept with the model's Ins /
    Del / Upd entries against the index built from scratch, every batch."""
    host = HostSession(spec=spec, batch_edges=be, threads=2)
    host.initial_keys()
    offs, adj, lab = host.graph()
    ql, qe = host.query_info()
    n = lab.size
    lists = [adj[int(offs[v]):int(offs[v + 1])].tolist() for v in range(n)]
    m = Model(n, lab.tolist(), lists, ql.tolist(), qe.reshape(-1, 2).tolist(), drop=drop)
    M = m.entries()
    mism, deaths = 0, 0
    rng = np.random.default_rng(seed)
    for pairs in deletion_stream(n, lists, host.update_stream().reshape(-1).tolist(), be, rng):
        lit0 = list(m.lit)
        new, gone_e, ins, dele, upd = m.batch(pairs)
        deaths += sum(1 for v in range(n) if lit0[v] & ~m.lit[v])
        M = (M | set(ins)) - set(dele)
        M |= set(upd)
        o2, a2 = _csr(m)
        if not np.array_equal(np.array(m.lit, dtype=np.uint32), O.li_fixpoint(o2, a2, lab, ql, qe)):
            mism += 1
        if M != m.entries():
            mism += 1
    return mism, deaths


DEL_STREAMS = [
    (synth_spec(9, 6, 3, seed=1, query_size=5, query_kind="dense", query_seed=1), 120),
    (synth_spec(9, 6, 2, seed=2, query_size=6, query_kind="sparse", query_seed=2), 100),
    (synth_spec(9, 8, 2, seed=3, query_size=5, query_kind="tree", query_seed=3), 150),
]


@pytest.mark.parametrize("i", range(len(DEL_STREAMS)))
def test_model_with_deletions(i):
    """Deletion streams: lit bits stay the fixpoint of the current graph
    and the index entries the deletions remove (over deleted edges, and of
    lit pairs that die) keep the main index equal to its from-scratch build;
    the streams kill pairs that were lit before their batch."""
    spec, be = DEL_STREAMS[i]
    mism, deaths = _run_deletions(spec, be)
    assert mism == 0
    assert deaths > 0


@pytest.mark.parametrize("drop", ["del_seed", "dedge_keys", "killed_keys"])
def test_model_without_a_deletion_rule_fails(drop):
    """Each deletion rule is needed: without the re-check of pairs that lost
    an edge, the lit set is not the fixpoint; without the entries of deleted
    edges or of killed pairs, the main index keeps stale entries."""
    assert sum(_run_deletions(spec, be, drop=drop)[0] for spec, be in DEL_STREAMS) > 0
