"""Binary CSR data input (a data path not ending in ".graph",
main_calig.cpp:93-101, index/calig.cpp:218-323): labels are one rand() draw
per vertex from the C library's initial state, the update stream is every
edge whose multiplicative hash of (min, max) is 0 mod 10, as (max, min)
pairs; the rest is the initial graph.  Checked against an independent
recomputation from the files (libc rand through ctypes, numpy hashing), and
the pipeline count against the closed form emb(G_end) - emb(G_0)."""
import ctypes
import ctypes.util

import numpy as np

from cgbdsm_amd.host import HostSession, generate_binary, synth_spec
from tests import oracle_lib as O


def _files(tmp_path, scale=9, labels=3, size=4, kind="dense", seed=4):
    spec = synth_spec(scale, 6, labels, seed=seed, query_size=size, query_kind=kind, query_seed=seed)
    qp, prefix = tmp_path / "q.graph", tmp_path / "g"
    generate_binary(spec, qp, prefix)
    return str(qp), str(prefix)


def _read(prefix):
    meta = open(prefix + ".meta.txt").read().split()
    n, m, nvcls = int(meta[0]), int(meta[1]), int(meta[8])
    offs = np.fromfile(prefix + ".vertex.bin", dtype=np.uint64)
    adj = np.fromfile(prefix + ".edge.bin", dtype=np.uint32)
    assert offs.size == n + 1 and adj.size == m == int(offs[-1])
    return n, nvcls, offs, adj


def test_binary_split_and_labels(tmp_path):
    qp, prefix = _files(tmp_path)
    n, nvcls, offs, adj = _read(prefix)
    host = HostSession(files=(qp, prefix, "unused"), threads=2)
    # labels: libc rand() from srand(1), kept only when the query uses them
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(1)
    draws = np.array([libc.rand() % nvcls for _ in range(n)], dtype=np.int64)
    ql, qe = host.query_info()
    want_lab = np.where(np.isin(draws, ql), draws, -1)
    g_offs, g_adj, g_lab = host.graph()
    assert np.array_equal(g_lab.astype(np.int64), want_lab)
    # stream: hash of (min, max) == 0 mod 10, as (max, min), each edge once
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(offs).astype(np.int64))
    dst = adj.astype(np.uint64)
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    key = (lo << np.uint64(32)) | hi
    with np.errstate(over="ignore"):
        h = key * np.uint64(0x9E3779B97F4A7C15)
    streamed = (h % np.uint64(10)) == 0
    one = src < dst
    want_stream = sorted(zip(hi[streamed & one].tolist(), lo[streamed & one].tolist()))
    got_stream = host.update_stream()
    assert sorted(map(tuple, got_stream.tolist())) == want_stream
    assert len(want_stream) > 0
    # initial graph: the other edges between vertices with a query label
    keep = ~streamed & one
    a, b = src[keep].astype(np.int64), dst[keep].astype(np.int64)
    ok = (want_lab[a] >= 0) & (want_lab[b] >= 0)
    want_edges = sorted(zip(a[ok].tolist(), b[ok].tolist()))
    gsrc = np.repeat(np.arange(g_offs.size - 1), np.diff(g_offs).astype(np.int64))
    gdst = g_adj[: int(g_offs[-1])].astype(np.int64)
    got_edges = sorted(zip(gsrc[gsrc < gdst].tolist(), gdst[gsrc < gdst].tolist()))
    assert got_edges == want_edges


def test_binary_pipeline_count_equals_closed_form(tmp_path):
    qp, prefix = _files(tmp_path, scale=9, labels=3, size=4, kind="dense", seed=6)
    host = HostSession(files=(qp, prefix, "unused"), batch_ints=200, threads=2)
    keys, dv, card, avg = host.initial_keys()
    plan = host.make_plan(card, avg)
    ql, qe = host.query_info()
    offs, adj, lab = host.graph()
    before = O.brute_count(offs, adj, lab, ql, qe)
    idx = O.Index(plan.ndir, 5)
    idx.set_dv(dv)
    idx.init(keys)
    total = 0
    for _ in range(host.num_batches):
        kb = host.next_batch()
        idx.set_dv(kb.dv)
        idx.insert(kb.ins)
        idx.remove(kb.dele)
        idx.rebuild_update(kb.upd)
        if kb.upd.size:
            c, _ = idx.match(plan, kb.edge_ok, masks=idx.masks(plan))
            total += c
            idx.insert(kb.upd)
    offs, adj, lab = host.graph()
    assert total == O.brute_count(offs, adj, lab, ql, qe) - before
    assert total > 0
