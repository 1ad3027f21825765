"""Semi-local read-out on the GPU against the textbook DP of every window."""
import numpy as np
import pytest

import lcs_amd
import oracle

pytestmark = pytest.mark.gpu


def test_string_substring_equals_dp(ctx):
    rng = np.random.default_rng(4)
    for t in range(6):
        m, n = int(rng.integers(1, 60)), int(rng.integers(1, 60))
        a = rng.integers(0, 3, m).astype(np.int32)
        b = rng.integers(0, 3, n).astype(np.int32)
        h, v, _ = ctx.comb(a, b)
        r2c = lcs_amd.from_strands(h, v)[0]
        qi, qj = np.triu_indices(n + 1)
        got = ctx.string_substring(r2c, m, n, qi, qj)
        exp = np.array([oracle.llcs_dp(a, b[i:j]) for i, j in zip(qi, qj)])
        assert np.array_equal(got, exp)


def test_dominance_counts_brute_force(ctx):
    rng = np.random.default_rng(5)
    p = rng.permutation(3000).astype(np.int32)
    qr = rng.integers(0, 3001, 500).astype(np.int32)
    qc = rng.integers(0, 3001, 500).astype(np.int32)
    got = ctx.dominance(p, qr, qc)
    exp = np.array([(p[r:] < c).sum() for r, c in zip(qr, qc)])
    assert np.array_equal(got, exp)


def test_windows_of_a_large_pair(ctx):
    w = lcs_amd.workload_single("mutated2^14", seed=9)
    h, v, llcs = ctx.comb(w.a, w.b)
    r2c = lcs_amd.from_strands(h, v)[0]
    n = w.b.size
    got = ctx.string_substring(r2c, w.a.size, n, [0, 100, 5000, 0], [n, 700, 9000, 0])
    assert got[0] == llcs and got[3] == 0
    assert got[1] == oracle.llcs_dp(w.a, w.b[100:700])
