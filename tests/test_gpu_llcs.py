"""The score-only bit-parallel LCS kernel (lcs_llcs*) against the textbook DP
(small pairs), the oracle's word-parallel CPU recurrence (large pairs) and
the seaweed comb's LCS read-out (the full 2^20 pair)."""
import numpy as np
import pytest

import lcs_amd
import oracle
from lcs_amd import hip as H

pytestmark = pytest.mark.gpu


def rnd(rng, m, n, sigma):
    return rng.integers(0, sigma, m).astype(np.int32), rng.integers(0, sigma, n).astype(np.int32)


@pytest.mark.parametrize("m,n", [(0, 0), (0, 5), (5, 0), (1, 1), (1, 300), (300, 1), (63, 64), (64, 65),
                                 (2047, 40), (2048, 2048), (2049, 100), (4097, 33), (33, 4097), (6000, 31),
                                 (31, 6000), (5000, 5000)])
def test_edge_shapes_equal_dp(ctx, m, n):
    rng = np.random.default_rng(m * 7919 + n)
    for sigma in (1, 2, 4, 26):
        a, b = rnd(rng, m, n, sigma)
        print(f"m={m} n={n} sigma={sigma}", flush=True)
        assert ctx.llcs(a, b, sigma) == oracle.llcs_dp(a, b)


def test_random_pairs_and_alphabets_equal_dp(ctx):
    rng = np.random.default_rng(17)
    for t in range(40):
        m, n = int(rng.integers(1, 3000)), int(rng.integers(1, 3000))
        sigma = int(rng.choice([1, 2, 3, 4, 8, 20, 26, 100, 200]))
        a, b = rnd(rng, m, n, sigma)
        if t % 3 == 0:   # long common runs
            b[: min(m, n) // 2] = a[: min(m, n) // 2]
        print(f"t={t} m={m} n={n} sigma={sigma}", flush=True)
        assert ctx.llcs(a, b, sigma) == oracle.llcs_dp(a, b)


@pytest.mark.parametrize("m,n,sigma", [(1 << 17, 1 << 17, 4), (1 << 18, 3000, 2), (3000, 1 << 18, 20),
                                       (100000, 70001, 200)])
def test_large_pairs_equal_oracle_bitpar(ctx, m, n, sigma):
    rng = np.random.default_rng(m + n + sigma)
    a = rng.integers(0, sigma, m).astype(np.int32)
    b = lcs_amd.gen_mutate(5, a, sigma, 0.1, 0.02, n) if m == n else rng.integers(0, sigma, n).astype(np.int32)
    print(f"m={m} n={n} sigma={sigma}", flush=True)
    assert ctx.llcs(a, b, sigma) == oracle.llcs_bitpar(a, b)


def test_equals_comb_readout(ctx):
    for name, seed in (("uniform4096", 3), ("mutated2^16", 4)):
        w = lcs_amd.workload_single(name, seed=seed)
        _, _, llcs = ctx.comb(w.a, w.b)
        print(name, flush=True)
        assert ctx.llcs(w.a, w.b, 4) == llcs


def test_full_size_pair_equals_comb(ctx):
    """m = n = 2^20 (the mutated pair): the bit-parallel length equals the
    seaweed permutation's LCS read-out."""
    w = lcs_amd.workload_single("mutated2^20", seed=1)
    _, v, llcs = ctx.comb(w.a, w.b)
    assert llcs == int((v < w.a.size).sum())
    assert ctx.llcs(w.a, w.b, 4) == llcs


def test_batch_of_uneven_pairs(ctx):
    rng = np.random.default_rng(23)
    pairs = []
    for p in range(40):
        m, n = int(rng.integers(0, 5000)), int(rng.integers(0, 5000))
        pairs.append(rnd(rng, m, n, 4))
    got = ctx.llcs_batch([p[0] for p in pairs], [p[1] for p in pairs], 4)
    exp = [oracle.llcs_bitpar(a, b) for a, b in pairs]
    assert list(got) == exp


def test_repeated_calls_do_not_see_stale_ring_words(ctx):
    """Two launches of the same shape back to back: the second must not read
    the first one's carry words."""
    rng = np.random.default_rng(99)
    for t in range(4):
        a, b = rnd(rng, 9000, 9000, 2 + t)
        assert ctx.llcs(a, b, 2 + t) == oracle.llcs_bitpar(a, b)


def test_device_entry(ctx):
    rng = np.random.default_rng(5)
    a, b = rnd(rng, 7000, 5000, 8)
    da, db, dout = H.DeviceArray.from_numpy(a), H.DeviceArray.from_numpy(b), H.DeviceArray(1)
    ctx.llcs_device(da.ptr.value, a.size, db.ptr.value, b.size, 8, dout.ptr.value)
    H.synchronize()
    assert int(dout.numpy()[0]) == oracle.llcs_bitpar(a, b)
    # a symbol outside [0, sigma) gives -1 on the device entry
    bad = b.copy()
    bad[17] = 8
    dbad = H.DeviceArray.from_numpy(bad)
    ctx.llcs_device(da.ptr.value, a.size, dbad.ptr.value, b.size, 8, dout.ptr.value)
    H.synchronize()
    assert int(dout.numpy()[0]) == -1


def test_symbols_outside_alphabet_are_rejected(ctx):
    a = np.array([0, 1, 2, 3], np.int32)
    for bad_a, bad_b in (([0, 1, 4], [1]), ([0, -1], [1]), ([0, 1], [1, 4]), ([1], [-5])):
        with pytest.raises(lcs_amd.LcsError, match="ERANGE"):
            ctx.llcs(np.array(bad_a, np.int32), np.array(bad_b, np.int32), 4)
    for sigma in (0, 201, -1):
        with pytest.raises(lcs_amd.LcsError, match="EINVAL"):
            ctx.llcs(a, a, sigma)
    assert ctx.llcs(a, a, 4) == 4
