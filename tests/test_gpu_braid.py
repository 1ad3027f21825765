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


# 2048 is the largest core kept wholly in one workgroup's LDS (keys,
# products and every depth's parents); larger cores run over the whole GPU
# (top depths device-wide, leaves of 2^11 points)
@pytest.mark.parametrize("k", [1, 2, 31, 32, 33, 100, 1000, 2047, 2048, 2049, 4096, 4097, 16384, 16385, 20000,
                               32768])
def test_core_sizes(ctx, k):
    rng = np.random.default_rng(k)
    p = rng.permutation(k).astype(np.int32)
    q = rng.permutation(k).astype(np.int32)
    assert np.array_equal(ctx.braid_multiply(p, q, k, False), oracle.staggered_multiply(p, q, k, False))


@pytest.mark.parametrize("k", [300, 2048])
def test_core_structured_permutations(ctx, k):
    """One-workgroup cores on permutations whose ant paths run straight for
    long stretches (identity, reversal, shifts, few swaps): the walks cut at
    row and column lines must still join up."""
    rng = np.random.default_rng(k)
    ident = np.arange(k, dtype=np.int32)
    near = ident.copy()
    for _ in range(20):
        i, j = rng.integers(0, k, 2)
        near[i], near[j] = near[j], near[i]
    cases = [(ident, ident[::-1].copy()), (ident[::-1].copy(), ident[::-1].copy()), (np.roll(ident, 37), near),
             (near, rng.permutation(k).astype(np.int32)), (ident[::-1].copy(), ident)]
    for p, q in cases:
        for glue in (False, True):
            assert np.array_equal(ctx.braid_multiply(p, q, k, glue), oracle.staggered_multiply(p, q, k, glue))


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


@pytest.mark.parametrize("k", [32769, 65536, 100000, 1 << 18])
def test_big_core_random_permutations(ctx, k):
    """Cores above 2^14 (the whole-GPU steady ant) against the oracle's
    staggered_multiply, both glue directions, with p and q wider than k."""
    rng = np.random.default_rng(k)
    print(f"big core k={k}", flush=True)
    for glue in (False, True):
        ps, qs = k + int(rng.integers(0, 2000)), k + int(rng.integers(0, 2000))
        p = rng.permutation(ps).astype(np.int32)
        q = rng.permutation(qs).astype(np.int32)
        got = ctx.braid_multiply(p, q, k, glue)
        assert np.array_equal(got, oracle.staggered_multiply(p, q, k, glue)), (k, glue)


def test_big_core_structured_permutations(ctx):
    """Permutations far from random (identity, reversal, shifts, few swaps)
    make the ant's path hug the box edges: every piece must still join."""
    k = 1 << 16
    rng = np.random.default_rng(3)
    ident = np.arange(k, dtype=np.int32)
    near = ident.copy()
    for _ in range(300):
        i, j = rng.integers(0, k, 2)
        near[i], near[j] = near[j], near[i]
    cases = [(ident, ident[::-1].copy()), (ident[::-1].copy(), ident[::-1].copy()), (np.roll(ident, 4321), near),
             (near, rng.permutation(k).astype(np.int32)), (ident[::-1].copy(), ident)]
    for p, q in cases:
        assert np.array_equal(ctx.braid_multiply(p, q, k, False), oracle.staggered_multiply(p, q, k, False))


def test_big_composition_m_2p20(ctx):
    """P(a, b1 b2) = P(a, b1) (x) P(a, b2) glued on k = m = 2^20, the strip
    permutations from GPU combs and the product from the GPU steady ant,
    against the oracle's comb of the whole pair."""
    w = lcs_amd.workload_single("mutated2^20", seed=2)
    a, b = w.a, w.b[: 3 * 8192]
    b1, b2 = b[:10000], b[10000:]
    print("composition m=2^20: combs", flush=True)
    P1 = ctx.comb_permutation(a, b1, lcs_amd.schedule(strips=1))[0]
    P2 = ctx.comb_permutation(a, b2, lcs_amd.schedule(strips=1))[0]
    print("composition m=2^20: product", flush=True)
    got = ctx.braid_multiply(P1, P2, a.size, False)
    ho, vo = oracle.comb(a, b, order="simd")
    assert np.array_equal(got, oracle.from_strands(ho, vo)[0])


@pytest.mark.parametrize("m,n,strips", [(1 << 16, 1 << 17, 4), (100000, 150000, 3), (1 << 20, 1 << 15, 2)])
def test_strips_above_32768_rows(ctx, m, n, strips):
    """Explicit column strips for pairs with m > 32768 (merged by the
    whole-GPU steady ant, k = m) give the oracle's strands."""
    rng = np.random.default_rng(m + n)
    a = rng.integers(0, 4, m).astype(np.int32)
    b = rng.integers(0, 4, n).astype(np.int32)
    print(f"strips m={m} n={n} strips={strips}", flush=True)
    h, v, _ = ctx.comb(a, b, lcs_amd.schedule(strips=strips))
    ho, vo = oracle.comb(a, b, order="simd")
    assert np.array_equal(h, ho) and np.array_equal(v, vo)
