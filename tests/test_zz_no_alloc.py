"""lcs_comb_device_noalloc (the entry point that allocates nothing: its
workspace is a static device array) against lcs_comb_device on the same
device buffers, bit for bit, on two calls in a row, the second with a
smaller input, for the register chain and the LDS band groups; sizes beyond
its workspace are refused."""
import ctypes

import numpy as np
import pytest

import lcs_amd
from lcs_amd import hip as H

pytestmark = pytest.mark.gpu


def _lib():
    L = lcs_amd.lib()
    vp = ctypes.c_void_p
    L.lcs_comb_device_noalloc.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, ctypes.POINTER(lcs_amd.Schedule),
                                          vp, vp, vp, vp]
    return L


def _run_both(ctx, L, w, sched):
    m, n = w.a.size, w.b.size
    da, db = H.DeviceArray.from_numpy(w.a), H.DeviceArray.from_numpy(w.b)
    h0, v0 = H.DeviceArray(m), H.DeviceArray(n)
    h1, v1 = H.DeviceArray(m), H.DeviceArray(n)
    st = H.DeviceArray.from_numpy(np.array([12345], np.int32))
    sp = ctypes.byref(sched) if sched is not None else None
    assert L.lcs_comb_device(ctx.handle, da.ptr, m, db.ptr, n, sp, h0.ptr, v0.ptr, None) == 0
    assert ctx.status() == lcs_amd.LCS_OK
    rc = L.lcs_comb_device_noalloc(da.ptr, m, db.ptr, n, sp, h1.ptr, v1.ptr, st.ptr, None)
    assert rc == 0
    H.synchronize()
    assert int(st.numpy()[0]) == 0
    assert np.array_equal(h0.numpy(), h1.numpy()) and np.array_equal(v0.numpy(), v1.numpy()), (m, n, sched)
    for d in (da, db, h0, v0, h1, v1, st):
        d.free()


@pytest.mark.parametrize("variant", [lcs_amd.VARIANT_CHAIN, lcs_amd.VARIANT_CHAIN_LDS])
def test_noalloc_equals_comb_device_twice(ctx, variant):
    L = _lib()
    big = lcs_amd.workload_single("mutated2^16", seed=11)
    small = lcs_amd.workload_batch(1, seed=12, lo=3000, hi=5000)[0]
    for w in (big, small):                      # the second call is smaller
        _run_both(ctx, L, w, lcs_amd.schedule(variant=variant))
    _run_both(ctx, L, small, None)              # default schedule
    _run_both(ctx, L, big, lcs_amd.schedule(tile_m=4))


def test_noalloc_empty_and_refused_sizes(ctx):
    L = _lib()
    a = H.DeviceArray.from_numpy(np.arange(5, dtype=np.int32))
    h = H.DeviceArray(5)
    assert L.lcs_comb_device_noalloc(a.ptr, 5, None, 0, None, h.ptr, None, None, None) == 0
    H.synchronize()
    assert list(h.numpy()) == [0, 1, 2, 3, 4]
    # b wider than the static ring: refused before any launch
    assert L.lcs_comb_device_noalloc(a.ptr, 5, a.ptr, (1 << 21) + 1, None, h.ptr, h.ptr, None, None) == -34
    # column strips need merges (more memory): refused
    assert L.lcs_comb_device_noalloc(a.ptr, 5, a.ptr, 5, ctypes.byref(lcs_amd.schedule(strips=2)), h.ptr, h.ptr,
                                     None, None) == -22
    a.free()
    h.free()
