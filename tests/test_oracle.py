"""Pins of the CPU oracle (runs on the CPU).

The reference holds no golden vectors or known-answer tests for this path,
so the oracle is pinned by brute force and closed forms: the textbook LCS
DP, the semi-local scores of every window, hand-derived tiny grids, the
composition law of the braid product against the comb, the definition of
the distance product, the reference's own PermutationMatrix compiled from
its source (oracle/_ref) and a big-integer restatement of its hash.
"""
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rnd(seed, m, n, sigma):
    rng = np.random.default_rng(seed)
    return rng.integers(0, sigma, m).astype(np.int32), rng.integers(0, sigma, n).astype(np.int32)


def test_tiny_known_answers():
    # a={1,2}, b={1}: row i=0 holds a[1]=2, row i=1 holds a[0]=1 (reversed).
    # Cell (1,0): 1==1 -> swap h=1,v=2 -> h[1]=2, v=1.  Cell (0,0): 2 vs 1
    # mismatch, h=0 > v=1? no -> keep.  So h=[0,2], v=[1].
    h, v = oracle.comb([1, 2], [1])
    assert list(h) == [0, 2] and list(v) == [1]
    # single matching cell swaps, single mismatching cell keeps
    assert [list(x) for x in oracle.comb([5], [5])] == [[1], [0]]
    assert [list(x) for x in oracle.comb([5], [6])] == [[0], [1]]


def test_no_match_is_identity():
    a = np.zeros(37, np.int32)
    b = np.ones(23, np.int32)
    h, v = oracle.comb(a, b)
    assert list(h) == list(range(37)) and list(v) == list(range(37, 60))


@pytest.mark.parametrize("order", ["rows", "antidiag", "mt"])
def test_llcs_equals_dp(order):
    for seed in range(60):
        rng = np.random.default_rng(seed)
        m, n, s = rng.integers(0, 40), rng.integers(0, 40), rng.integers(1, 6)
        a, b = rnd(seed, m, n, s)
        h, v = oracle.comb(a, b, order=order, threads=3)
        assert oracle.llcs_from_strands(v, m) == oracle.llcs_dp(a, b)
        r2c, c2r = oracle.from_strands(h, v)
        assert sorted(r2c) == list(range(m + n))


def test_equal_strings_and_orders_agree():
    a, b = rnd(3, 500, 700, 4)
    assert oracle.llcs_from_strands(oracle.comb(a, a)[1], 500) == 500
    r = oracle.comb(a, b, "rows")
    for order in ("antidiag", "mt"):
        o = oracle.comb(a, b, order, threads=5)
        assert np.array_equal(r[0], o[0]) and np.array_equal(r[1], o[1])


SIMD_ISAS = [i for i in (oracle.ISA_AVX2, oracle.ISA_AVX512) if oracle.simd_isa_available(i)]


@pytest.mark.parametrize("isa", SIMD_ISAS)
def test_simd_comb_equals_rows(isa):
    """The vectorised stripe comb (the CPU baseline and the checker of the
    full-size pairs) against the sequential comb: stripe edges (8 or 16 lanes
    of 8 rows), column-block edges (2048), empty grids, negative and 32-bit
    symbols, all-match and no-match, 1..8 threads (the pipeline over threads)."""
    rng = np.random.default_rng(11)
    W = 16 if isa == oracle.ISA_AVX512 else 8
    edges = [0, 1, 7, 8, 9, W - 1, W, W + 1, 8 * W - 1, 8 * W, 8 * W + 1, 2047, 2048, 2049]
    cases = [(m, n) for m in edges for n in (0, 1, 5, W + 1, 2049)]
    cases += [(int(rng.integers(0, 3000)), int(rng.integers(0, 6000))) for _ in range(40)]
    for k, (m, n) in enumerate(cases):
        sigma = int(rng.integers(1, 6))
        a, b = rnd(1000 + k, m, n, sigma)
        if k % 7 == 3:   # negative and 32-bit symbols
            a = (a.astype(np.int64) * 0x3fffffff - 0x40000000).astype(np.int32)
            b = (b.astype(np.int64) * 0x3fffffff - 0x40000000).astype(np.int32)
        ho, vo = oracle.comb(a, b)
        for th in (1, 3, 8):
            h, v = oracle.comb(a, b, order="simd", threads=th, isa=isa)
            assert np.array_equal(h, ho) and np.array_equal(v, vo), (m, n, sigma, th)
    # all-match and no-match grids
    z, o = np.zeros(333, np.int32), np.ones(4100, np.int32)
    for a, b in ((z, z[:100]), (o, o), (z, o)):
        ho, vo = oracle.comb(a, b)
        h, v = oracle.comb(a, b, order="simd", threads=4, isa=isa)
        assert np.array_equal(h, ho) and np.array_equal(v, vo)


def test_semilocal_windows_equal_dp():
    for seed in range(25):
        rng = np.random.default_rng(100 + seed)
        m, n, s = rng.integers(1, 25), rng.integers(1, 25), rng.integers(1, 4)
        a, b = rnd(seed, m, n, s)
        r2c = oracle.permutation(a, b)
        H = oracle.string_substring(r2c, m, n)
        for i in range(n + 1):
            for j in range(i, n + 1):
                assert H[i, j] == oracle.llcs_dp(a, b[i:j])


def test_semilocal_quadrants_equal_dp():
    """Every query of the four quadrants against the DP of its substrings,
    empty strings included; dominance against the definition; -1 outside."""
    for seed in range(30):
        rng = np.random.default_rng(300 + seed)
        m, n, s = int(rng.integers(0, 14)), int(rng.integers(0, 14)), int(rng.integers(1, 4))
        a, b = rnd(seed + 7, m, n, s)
        r2c = oracle.permutation(a, b)
        L = oracle.llcs_dp
        qs = {0: [(i, j) for i in range(n + 1) for j in range(i, n + 1)],
              1: [(k, e) for k in range(m + 1) for e in range(k, m + 1)],
              2: [(p, i) for p in range(m + 1) for i in range(n + 1)],
              3: [(k, j) for k in range(m + 1) for j in range(n + 1)]}
        want = {0: lambda x, y: L(a, b[x:y]), 1: lambda x, y: L(a[x:y], b),
                2: lambda x, y: L(a[:x], b[y:]), 3: lambda x, y: L(a[x:], b[:y])}
        for kind, q in qs.items():
            q0, q1 = np.array(q, np.int32).T
            got = oracle.semilocal(r2c, m, n, kind, q0, q1)
            assert list(got) == [want[kind](x, y) for x, y in q], (seed, kind)
        r0 = rng.integers(-2, m + n + 2, 40).astype(np.int32)
        c0 = rng.integers(-2, m + n + 2, 40).astype(np.int32)
        dom = oracle.semilocal(r2c, m, n, 4, r0, c0)
        assert list(dom) == [int((r2c[max(int(x), 0):] < y).sum()) for x, y in zip(r0, c0)]
    r2c = oracle.permutation([1, 2], [2, 1, 2])
    bad = oracle.semilocal(r2c, 2, 3, 0, [2, -1, 0], [1, 3, 4])
    assert list(bad) == [-1, -1, -1]
    assert list(oracle.semilocal(r2c, 2, 3, 1, [0, 1], [3, 1])) == [-1, 0]


def test_llcs_bitpar_equals_dp():
    rng = np.random.default_rng(41)
    sizes = [0, 1, 2, 63, 64, 65, 127, 128, 129]
    for m in sizes:
        for n in (0, 1, 5, 64, 130):
            a, b = rng.integers(0, 3, m), rng.integers(0, 3, n)
            assert oracle.llcs_bitpar(a, b) == oracle.llcs_dp(a, b), (m, n)
    for seed in range(40):
        rng = np.random.default_rng(900 + seed)
        m, n, s = int(rng.integers(1, 400)), int(rng.integers(1, 400)), int(rng.integers(1, 31))
        lo = int(rng.integers(-2 ** 31, 2 ** 31 - 64))
        a = rng.integers(lo, lo + s, m).astype(np.int32)
        b = rng.integers(lo, lo + s + 2, n).astype(np.int32)
        assert oracle.llcs_bitpar(a, b) == oracle.llcs_dp(a, b), seed
    assert oracle.llcs_bitpar([1, 2, 3, 4], [2, 4]) == 2


def test_dp_hand_cases():
    assert oracle.llcs_dp([1, 2, 3, 4], [2, 4]) == 2
    assert oracle.llcs_dp([1, 2, 3], [3, 2, 1]) == 1
    assert oracle.llcs_dp([], [1, 2]) == 0
    assert oracle.llcs_dp([0, 1, 0, 1, 1], [1, 0, 1, 1, 0]) == 4


def test_fixup_quadrants_equal_whole():
    a, b = rnd(9, 61, 83, 3)
    m, n = 61, 83
    ap = a[::-1].copy()
    h = np.arange(m, dtype=np.int32)
    v = (m + np.arange(n)).astype(np.int32)
    # dependency order: rows with high i first (v flows to low i), then left to right
    oracle.fixup(ap, b, h, v, 20, m - 20, 0, 40)
    oracle.fixup(ap, b, h, v, 0, 20, 0, 40)
    oracle.fixup(ap, b, h, v, 0, m, 40, n - 40)
    ho, vo = oracle.comb(a, b)
    assert np.array_equal(h, ho) and np.array_equal(v, vo)


def _hash_bigint(c2r):
    R, M = 4294967279, 4294967291
    res = 0
    for x in c2r:
        t = (R * res + int(x)) % (1 << 64)
        if t >= 1 << 63:
            t -= 1 << 64
        # C remainder truncates toward zero
        res = abs(t) % M * (1 if t >= 0 else -1)
    return res


def test_hash_wraps_like_int64():
    a, b = rnd(4, 300, 400, 2)
    h, v = oracle.comb(a, b)
    c2r = oracle.from_strands(h, v)[1]
    assert oracle.perm_hash(c2r) == _hash_bigint(c2r)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_from_strands_and_hash_match_reference_build():
    for seed in range(20):
        rng = np.random.default_rng(seed)
        m, n = rng.integers(0, 400), rng.integers(0, 400)
        a, b = rnd(seed, m, n, 2)
        h, v = oracle.comb(a, b)
        r2c, c2r = oracle.from_strands(h, v)
        rr, rc, hv = oracle.ref_from_strands_hash(h, v)
        assert np.array_equal(rr, r2c) and np.array_equal(rc, c2r) and hv == oracle.perm_hash(c2r)


def test_composition_law_both_glues():
    rng = np.random.default_rng(77)
    for t in range(150):
        s = rng.integers(1, 5)
        m, n1, n2 = rng.integers(1, 25), rng.integers(0, 25), rng.integers(0, 25)
        a, b1 = rnd(t, m, n1, s)
        _, b2 = rnd(t + 1000, 0, n2, s)
        P = oracle.permutation(a, np.concatenate([b1, b2]))
        R = oracle.staggered_multiply(oracle.permutation(a, b1), oracle.permutation(a, b2), m, False)
        assert np.array_equal(P, R)
        m1, m2, n = rng.integers(0, 25), rng.integers(0, 25), rng.integers(1, 25)
        a1, b = rnd(t + 2000, m1, n, s)
        a2, _ = rnd(t + 3000, m2, 0, s)
        P = oracle.permutation(np.concatenate([a1, a2]), b)
        R = oracle.staggered_multiply(oracle.permutation(a1, b), oracle.permutation(a2, b), n, True)
        assert np.array_equal(P, R)


def test_steady_ant_equals_distance_product_definition():
    rng = np.random.default_rng(5)
    for t in range(300):
        k = int(rng.integers(1, 10))
        p = rng.permutation(k).astype(np.int32)
        q = rng.permutation(k).astype(np.int32)
        assert np.array_equal(oracle.steady_ant(p, q), oracle.sticky_bruteforce(p, q))


def test_golden_fixtures():
    """Committed fixtures (tests/golden/) reproduce."""
    path = os.path.join(GOLDEN, "comb_fixtures.json")
    with open(path) as f:
        fx = json.load(f)
    for case in fx["cases"]:
        a = np.array(case["a"], np.int32)
        b = np.array(case["b"], np.int32)
        h, v = oracle.comb(a, b)
        assert list(h) == case["h"] and list(v) == case["v"]
        assert oracle.perm_hash(oracle.from_strands(h, v)[1]) == case["hash"]


def test_golden_large_digests():
    import hashlib
    import lcs_amd
    with open(os.path.join(GOLDEN, "comb_fixtures.json")) as f:
        fx = json.load(f)
    for c in fx["large"]:
        if c["kind"] == "binary":
            a, b = lcs_amd.gen_binary(c["seed_a"], c["m"]), lcs_amd.gen_binary(c["seed_b"], c["n"])
        else:
            a, b = lcs_amd.gen_uniform(c["seed_a"], c["m"], 4), lcs_amd.gen_uniform(c["seed_b"], c["n"], 4)
        h, v = oracle.comb(a, b, "mt", threads=4)
        assert hashlib.sha256(np.concatenate([h, v]).astype("<i4").tobytes()).hexdigest() == c["sha256"]
        assert oracle.perm_hash(oracle.from_strands(h, v)[1]) == c["hash"]
