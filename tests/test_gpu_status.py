"""Asynchronous error reporting and the shared workspace across streams.

A device-side wait that gives up is forced from the host by a zero wait
bound (lcs_ctx_set_timeout): every wave still exits normally, the launch's
outputs are undefined, and lcs_ctx_status reports LCS_ETIMEDOUT once."""
import numpy as np
import pytest

import lcs_amd
import oracle
from lcs_amd import hip as H

pytestmark = pytest.mark.gpu


def _device_pair(w):
    return (H.DeviceArray.from_numpy(w.a), H.DeviceArray.from_numpy(w.b), H.DeviceArray(w.a.size),
            H.DeviceArray(w.b.size))


def test_async_abort_is_reported_by_status(ctx):
    L = lcs_amd.lib()
    w = lcs_amd.workload_single("mutated2^16", seed=4)   # 2^16 rows: 1024 bands, band g waits on g-1
    m, n = w.a.size, w.b.size
    da, db, dh, dv = _device_pair(w)
    assert ctx.status() == lcs_amd.LCS_OK
    ctx.set_timeout(0)
    try:
        assert L.lcs_comb_device(ctx.handle, da.ptr, m, db.ptr, n, None, dh.ptr, dv.ptr, None) == 0
        assert ctx.status() == lcs_amd.LCS_ETIMEDOUT
        assert ctx.status() == lcs_amd.LCS_OK            # the report is cleared
        # the bit-parallel LCS length chain reports the same way
        dout = H.DeviceArray(1)
        ctx.llcs_device(da.ptr.value, m, db.ptr.value, n, 4, dout.ptr.value)
        assert ctx.status() == lcs_amd.LCS_ETIMEDOUT
        # a synchronous call reports its own abort and leaves no stale report
        with pytest.raises(lcs_amd.LcsError) as e:
            ctx.comb(w.a, w.b)
        assert e.value.rc == lcs_amd.LCS_ETIMEDOUT
        assert ctx.status() == lcs_amd.LCS_OK
    finally:
        ctx.set_timeout(lcs_amd.DEFAULT_TIMEOUT_TICKS)
    # the context is healthy again: the next launch clears its control words
    assert L.lcs_comb_device(ctx.handle, da.ptr, m, db.ptr, n, None, dh.ptr, dv.ptr, None) == 0
    assert ctx.status() == lcs_amd.LCS_OK
    ho, vo = oracle.comb(w.a, w.b, order="simd")
    assert np.array_equal(dh.numpy(), ho) and np.array_equal(dv.numpy(), vo)


def test_device_calls_on_different_streams_are_serialised(ctx):
    """Calls through one context share its ring, descriptors and control
    words: a call on another stream waits for the previous call's kernels,
    so back-to-back calls on two streams both come out right."""
    L = lcs_amd.lib()
    s1, s2 = H.Stream(), H.Stream()
    try:
        works = [lcs_amd.workload_single("mutated2^16", seed=5), lcs_amd.workload_single("uniform4096", seed=6),
                 lcs_amd.workload_single("mutated2^15", seed=7)]
        bufs = [_device_pair(w) for w in works]
        for (w, (da, db, dh, dv), s) in zip(works, bufs, (s1, s2, s1)):
            rc = L.lcs_comb_device(ctx.handle, da.ptr, w.a.size, db.ptr, w.b.size, None, dh.ptr, dv.ptr, s.ptr)
            assert rc == 0
        assert ctx.status() == lcs_amd.LCS_OK
        s1.synchronize()
        s2.synchronize()
        for w, (da, db, dh, dv) in zip(works, bufs):
            ho, vo = oracle.comb(w.a, w.b, order="simd")
            assert np.array_equal(dh.numpy(), ho) and np.array_equal(dv.numpy(), vo), (w.a.size, w.b.size)
    finally:
        s1.close()
        s2.close()
