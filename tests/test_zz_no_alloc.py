"""The sort entry point that allocates nothing (arsim_sort_pairs_u32_noalloc) gives the same keys,
values and pointer exchange as arsim_sort_pairs_u32 with a caller workspace.

Two calls in a row share the library's static scratch array: the first at the largest size the
array holds, the second smaller and over a different bit range, so that anything the first call
left behind (digit histograms, tile counters, look-back words) would show up in the second.  Both
are compared bit for bit with the workspace entry point and with numpy's stable argsort.
"""
import ctypes as C

import numpy as np
import pytest

from diploma_thesis_amd import capi
from gpu_util import Dev, sync, zeros

pytestmark = pytest.mark.gpu
L = capi.lib
MAX_ITEMS = 1 << 22


def _sort(keys, vals, begin, end, noalloc):
    n = keys.size
    dk, dv = Dev(keys), Dev(vals)
    ak, av = zeros(n, np.uint32), zeros(n, np.uint32)
    pk, pv, pak, pav = (C.c_void_p(x.p.value) for x in (dk, dv, ak, av))
    if noalloc:
        rc = L.arsim_sort_pairs_u32_noalloc(C.byref(pk), C.byref(pv), C.byref(pak), C.byref(pav), n, begin, end,
                                            None)
    else:
        ws_bytes = L.arsim_sort_workspace_bytes(n, 4)
        ws = Dev(shape=(max(ws_bytes, 4),), dtype=np.uint8)
        rc = L.arsim_sort_pairs_u32(C.byref(pk), C.byref(pv), C.byref(pak), C.byref(pav), n, begin, end, ws.p,
                                    ws_bytes, None)
    capi.check(rc, "noalloc" if noalloc else "workspace")
    sync()
    assert {pk.value, pak.value} == {dk.p.value, ak.p.value}
    assert {pv.value, pav.value} == {dv.p.value, av.p.value}
    # where *keys / *values point after the call: the sorted data (cub::DoubleBuffer semantics)
    in_first = (pk.value == dk.p.value, pv.value == dv.p.value)
    out_k = dk.get() if in_first[0] else ak.get()
    out_v = dv.get() if in_first[1] else av.get()
    return out_k, out_v, in_first


def test_noalloc_sort_equals_workspace_sort_twice():
    rng = np.random.default_rng(22)
    cases = [(MAX_ITEMS, rng.integers(0, 1 << 24, MAX_ITEMS, dtype=np.uint32), 0, 24),
             (70001, rng.integers(0, 1 << 32, 70001, dtype=np.uint64).astype(np.uint32), 8, 20)]
    for n, keys, b, e in cases:
        print(f"noalloc sort: n={n} bits [{b},{e})", flush=True)
        vals = rng.integers(0, 1 << 30, n, dtype=np.uint32)
        nk, nv, nside = _sort(keys, vals, b, e, noalloc=True)
        wk, wv, wside = _sort(keys, vals, b, e, noalloc=False)
        assert nside == wside
        assert np.array_equal(nk, wk) and np.array_equal(nv, wv), n
        digits = (keys.astype(np.uint64) >> np.uint64(b)) & np.uint64((1 << (e - b)) - 1)
        order = np.argsort(digits, kind="stable")
        assert np.array_equal(nk, keys[order]) and np.array_equal(nv, vals[order]), n


def test_noalloc_sort_refuses_more_items_than_the_scratch_holds():
    d = zeros(16, np.uint32)
    p = [C.c_void_p(d.p.value) for _ in range(4)]
    before = [x.value for x in p]
    rc = L.arsim_sort_pairs_u32_noalloc(C.byref(p[0]), C.byref(p[1]), C.byref(p[2]), C.byref(p[3]), MAX_ITEMS + 1,
                                        0, 32, None)
    assert rc == capi.ARSIM_ERR_WORKSPACE
    assert b"static scratch" in L.arsim_last_error()
    assert [x.value for x in p] == before
