"""Parity of the GPU steady ant (lcs_braid_multiply) with the oracle's
restatement of staggered_multiply, and the composition law against the
comb at sizes the oracle cannot merge quickly."""
import numpy as np
import pytest

import lcs_amd
import oracle

pytestmark = pytest.mark.gpu


def test_core_random_permutations(ctx):
    rng = np.random.default_rng(1)
    for t in range(200):
        k = int(rng.integers(1, 70))
        p = rng.permutation(k).astype(np.int32)
        q = rng.permutation(k).astype(np.int32)
        for glue in (False, True):
            got = ctx.braid_multiply(p, q, k, glue)
            exp = oracle.staggered_multiply(p, q, k, glue)
            assert np.array_equal(got, exp), (k, glue)


# 16384 is the largest core kept wholly in LDS; 20000 and 32768 take the
# global-memory form of the same kernel (16-bit entries up to K = 32768)
@pytest.mark.parametrize("k", [1, 2, 31, 32, 33, 100, 1000, 4097, 16384, 16385, 20000, 32768])
def test_core_sizes(ctx, k):
    rng = np.random.default_rng(k)
    p = rng.permutation(k).astype(np.int32)
    q = rng.permutation(k).astype(np.int32)
    assert np.array_equal(ctx.braid_multiply(p, q, k, False), oracle.staggered_multiply(p, q, k, False))


def test_glued_products_of_combs(ctx):
    rng = np.random.default_rng(7)
    for t in range(40):
        s = int(rng.integers(2, 5))
        m, n1, n2 = (int(x) for x in rng.integers(1, 300, 3))
        a = rng.integers(0, s, m).astype(np.int32)
        b1 = rng.integers(0, s, n1).astype(np.int32)
        b2 = rng.integers(0, s, n2).astype(np.int32)
        P = oracle.permutation(a, np.concatenate([b1, b2]))
        got = ctx.braid_multiply(oracle.permutation(a, b1), oracle.permutation(a, b2), m, False)
        assert np.array_equal(got, P)
        a2 = rng.integers(0, s, int(rng.integers(1, 300))).astype(np.int32)
        P = oracle.permutation(np.concatenate([a, a2]), b1)
        got = ctx.braid_multiply(oracle.permutation(a, b1), oracle.permutation(a2, b1), n1, True)
        assert np.array_equal(got, P)


def test_large_b_split_composition(ctx):
    # P(a, b1 b2) from two GPU combs and one GPU merge, k = m = 8192
    w = lcs_amd.workload_single("mutated2^13", seed=3)
    a, b = w.a, w.b
    b1, b2 = b[:3000], b[3000:]
    h1, v1, _ = ctx.comb(a, b1)
    h2, v2, _ = ctx.comb(a, b2)
    h, v, _ = ctx.comb(a, b)
    P1 = lcs_amd.from_strands(h1, v1)[0]
    P2 = lcs_amd.from_strands(h2, v2)[0]
    P = lcs_amd.from_strands(h, v)[0]
    assert np.array_equal(ctx.braid_multiply(P1, P2, a.size, False), P)


def test_k_zero_and_bad_input(ctx):
    p = np.array([1, 0], np.int32)
    q = np.array([0, 2, 1], np.int32)
    assert np.array_equal(ctx.braid_multiply(p, q, 0, False), oracle.staggered_multiply(p, q, 0, False))
    assert np.array_equal(ctx.braid_multiply(p, q, 0, True), oracle.staggered_multiply(p, q, 0, True))
    with pytest.raises(lcs_amd.LcsError) as e:
        ctx.braid_multiply(np.array([0, 0], np.int32), q, 1, False)
    assert e.value.rc == -22
