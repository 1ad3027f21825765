"""Continuous batching: requests arriving while others run are admitted as
slots and memory free up, overlap on the device, and each result equals the
one the same pair gets alone (bit-exact: integer output)."""
import time

import numpy as np
import pytest

import lcs_amd
import oracle

pytestmark = pytest.mark.gpu


def test_results_alone_equal_results_in_the_mix(ctx):
    pairs = lcs_amd.workload_batch(48, seed=21, lo=512, hi=20000, alphabet=(4, 20))
    srv = lcs_amd.Server(ctx, hbm_budget=64 << 20, max_inflight=3)
    rng = np.random.default_rng(0)
    ids = []
    for p in pairs:                     # Poisson arrivals while earlier ones run
        ids.append(srv.submit(p.a, p.b))
        time.sleep(rng.exponential(0.002))
    outs = [srv.wait(i, timeout_ms=60000) for i in ids]
    st = srv.stats()
    assert st["completed"] == len(pairs) and st["status"] == 0
    assert st["batches"] > 1            # admitted over several micro-batches
    assert st["max_inflight"] >= 2      # micro-batches overlapped on the device
    for p, (h, v) in zip(pairs, outs):
        h1, v1, _ = ctx.comb(p.a, p.b)  # the same pair alone
        assert np.array_equal(h, h1) and np.array_equal(v, v1)
    for p, (h, v) in list(zip(pairs, outs))[:8]:
        ho, vo = oracle.comb(p.a, p.b)
        assert np.array_equal(h, ho) and np.array_equal(v, vo)
    for i in ids:
        q, t = srv.times(i)
        assert 0 <= q <= t
        srv.release(i)
    srv.close()


def test_budget_admission_and_errors(ctx):
    srv = lcs_amd.Server(ctx, hbm_budget=4 << 20, max_inflight=2)
    big = np.zeros(600_000, np.int32)
    with pytest.raises(lcs_amd.LcsError) as e:   # could never fit the budget
        srv.submit(big, big)
    assert e.value.rc == -12
    a = np.arange(3000, dtype=np.int32) % 7
    ids = [srv.submit(a, a[::-1].copy()) for _ in range(20)]   # more than fits at once
    outs = [srv.wait(i, timeout_ms=60000) for i in ids]
    h1, v1, _ = ctx.comb(a, a[::-1].copy())
    assert all(np.array_equal(h, h1) and np.array_equal(v, v1) for h, v in outs)
    assert lcs_amd.lib().lcs_server_cancel(srv._h, 10**9) == -2   # unknown request id: LCS_ENOENT
    srv.close()


def test_cost_model_swapped_while_requests_run(ctx):
    """The model handed to the server is copied: the worker builds its own
    on the server's context, so the caller may detach it, attach another
    or destroy its own while requests run, and every result stays equal to
    the pair combed alone."""
    from lcs_amd import tuning
    mlp = lcs_amd.Mlp(ctx, n_in=4, hidden=16, n_out=len(tuning.CANDIDATES), seed=3)
    pairs = lcs_amd.workload_batch(40, seed=23, lo=512, hi=12000, alphabet=(4, 20))
    feats = tuning.features([p.a.size for p in pairs], [p.b.size for p in pairs], [p.alphabet for p in pairs],
                            [p.rate for p in pairs])
    srv = lcs_amd.Server(ctx, hbm_budget=64 << 20, max_inflight=3)
    try:
        srv.set_cost_model(mlp, tuning.CANDIDATES)
        ids = []
        for i, (p, f) in enumerate(zip(pairs, feats)):
            ids.append(srv.submit(p.a, p.b, None, f))
            if i == 10:
                srv.set_cost_model(None, [])               # detach mid-stream
            if i == 20:
                srv.set_cost_model(mlp, tuning.CANDIDATES)  # attach again
            if i == 30:
                mlp.close()                                 # the caller's model goes away
        outs = [srv.wait(i, timeout_ms=60000) for i in ids]
        assert srv.stats()["status"] == 0
    finally:
        srv.close()
        mlp.close()
    for p, (h, v) in zip(pairs, outs):
        h1, v1, _ = ctx.comb(p.a, p.b)
        assert np.array_equal(h, h1) and np.array_equal(v, v1)
