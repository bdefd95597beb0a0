"""cgbdsm_index_export_static (the export that allocates nothing on the
device) returns exactly what cgbdsm_index_export returns, on two calls in a
row (the second on a smaller index), for the main and the update index, and
rejects an index beyond its static scratch and a too-small output buffer."""
import ctypes as C

import numpy as np
import pytest

from cgbdsm_amd.engine import CgbdsmError, Engine, _check
from cgbdsm_amd.host import HostSession, synth_spec

pytestmark = pytest.mark.gpu

STATIC_WORDS = 4 << 20


def _static_export(q, which, cap=None):
    n = C.c_uint64()
    if cap is None:
        _check(q.L.cgbdsm_index_export(q.h, which, None, 0, C.byref(n)))
        cap = n.value
    out = np.zeros(max(cap, 1), dtype=np.uint64)
    _check(q.L.cgbdsm_index_export_static(q.h, which, out.ctypes.data, cap, C.byref(n)))
    return out[: n.value]


def _query_after_one_batch(eng, scale, deg, seed, layout):
    spec = synth_spec(scale, deg, 4, seed=seed, query_size=5, query_kind="dense", query_seed=seed)
    host = HostSession(spec=spec, batch_edges=400, threads=4)
    keys, dv, card, avg = host.initial_keys()
    q = eng.query(host.query_tables)
    q.set_layout(layout)
    q.index_init(keys, dv)
    q.set_plan(host.make_plan(card, avg))
    q.stage(host.next_batch())
    q.transfer_update_all()
    q.transfer_update_index()
    return q


@pytest.mark.parametrize("layout", ["gamma", "gpma"])
def test_static_export_equals_export_twice(layout):
    print(f"case static export {layout}", flush=True)
    eng = Engine(0)
    big = _query_after_one_batch(eng, 12, 10, 3, layout)
    small = _query_after_one_batch(eng, 9, 6, 4, layout)
    sizes = []
    for q in (big, small):  # second call on a smaller index: nothing left over
        for which in (0, 1):
            a = q.export(which)
            b = _static_export(q, which)
            assert a.size > 0
            assert np.array_equal(a, b), (which, a.size, b.size)
            sizes.append(a.size)
        bound = sizes[-2] + sizes[-1]  # M after the commit holds at most M + U entries
        q.build_local_index()
        q.matching()
        q.batch_update_global_index()
        # M after the commit; under GPMA the flat lists are stale here and
        # the static export refreshes them itself, without pool growth.
        b = _static_export(q, 0, cap=bound)
        assert np.array_equal(q.export(0), b)
    assert sizes[2] < sizes[0]
    big.close()
    small.close()
    eng.close()


def test_static_export_limits():
    print("case static export limits", flush=True)
    eng = Engine(0)
    q = _query_after_one_batch(eng, 9, 6, 5, "gamma")
    n = q.export(0).size
    out = np.zeros(max(n - 1, 1), dtype=np.uint64)
    got = C.c_uint64()
    rc = q.L.cgbdsm_index_export_static(q.h, 0, out.ctypes.data, n - 1, C.byref(got))
    assert rc == 1 and got.value == n  # INVALID, with the size reported
    q.close()
    # An index whose rows and entries exceed the static scratch is refused.
    spec = synth_spec(17, 8, 1, seed=6, query_size=4, query_kind="tree", query_seed=6)
    host = HostSession(spec=spec, batch_edges=2000, threads=8)
    keys, dv, card, avg = host.initial_keys()
    q2 = eng.query(host.query_tables)
    q2.index_init(keys, dv)
    n2 = q2.export(0).size
    assert dv * q2.ndir + n2 > STATIC_WORDS
    with pytest.raises(CgbdsmError) as e:
        _static_export(q2, 0, cap=n2)
    assert e.value.code == 4  # LIMIT
    # The same context still exports a small index afterwards.
    q3 = _query_after_one_batch(eng, 9, 6, 7, "gamma")
    assert np.array_equal(q3.export(1), _static_export(q3, 1))
    q2.close()
    q3.close()
    eng.close()
