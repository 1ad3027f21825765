"""Semi-local read-out on the GPU: the four quadrants and raw dominance
counts, by the direct-count kernel and by the wavelet-matrix index, against
the textbook DP of every window (small pairs) and the oracle's direct count
(large permutations)."""
import numpy as np
import pytest

import lcs_amd
import oracle
from lcs_amd import hip as H

pytestmark = pytest.mark.gpu

KINDS = (lcs_amd.SEMI_STRING_SUBSTRING, lcs_amd.SEMI_SUBSTRING_STRING, lcs_amd.SEMI_PREFIX_SUFFIX,
         lcs_amd.SEMI_SUFFIX_PREFIX)


def all_queries(kind, m, n):
    if kind == lcs_amd.SEMI_STRING_SUBSTRING:
        q = [(i, j) for i in range(n + 1) for j in range(i, n + 1)]
    elif kind == lcs_amd.SEMI_SUBSTRING_STRING:
        q = [(k, e) for k in range(m + 1) for e in range(k, m + 1)]
    else:
        q = [(x, y) for x in range(m + 1) for y in range(n + 1)]
    return np.array(q, np.int32).reshape(-1, 2).T


def dp(kind, a, b, x, y):
    return {lcs_amd.SEMI_STRING_SUBSTRING: lambda: oracle.llcs_dp(a, b[x:y]),
            lcs_amd.SEMI_SUBSTRING_STRING: lambda: oracle.llcs_dp(a[x:y], b),
            lcs_amd.SEMI_PREFIX_SUFFIX: lambda: oracle.llcs_dp(a[:x], b[y:]),
            lcs_amd.SEMI_SUFFIX_PREFIX: lambda: oracle.llcs_dp(a[x:], b[:y])}[kind]()


def random_queries(rng, kind, m, n, nq):
    if kind == lcs_amd.SEMI_STRING_SUBSTRING:
        x, y = rng.integers(0, n + 1, nq), rng.integers(0, n + 1, nq)
        return np.minimum(x, y).astype(np.int32), np.maximum(x, y).astype(np.int32)
    if kind == lcs_amd.SEMI_SUBSTRING_STRING:
        x, y = rng.integers(0, m + 1, nq), rng.integers(0, m + 1, nq)
        return np.minimum(x, y).astype(np.int32), np.maximum(x, y).astype(np.int32)
    if kind == lcs_amd.SEMI_DOMINANCE:
        return (rng.integers(-2, m + n + 2, nq).astype(np.int32), rng.integers(-2, m + n + 2, nq).astype(np.int32))
    return rng.integers(0, m + 1, nq).astype(np.int32), rng.integers(0, n + 1, nq).astype(np.int32)


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


def test_four_quadrants_equal_dp_scan_and_index(ctx):
    """Every query of every quadrant, empty strings included, by both the
    direct count and the index."""
    rng = np.random.default_rng(12)
    for t in range(14):
        m, n = int(rng.integers(0, 24)), int(rng.integers(0, 24))
        if t == 0:
            m, n = 0, 0
        sig = int(rng.integers(1, 5))
        a = rng.integers(0, sig, m).astype(np.int32)
        b = rng.integers(0, sig, n).astype(np.int32)
        h, v, _ = ctx.comb(a, b)
        r2c = lcs_amd.from_strands(h, v)[0]
        idx = lcs_amd.DominanceIndex(ctx, r2c)
        for kind in KINDS:
            q0, q1 = all_queries(kind, m, n)
            exp = np.array([dp(kind, a, b, x, y) for x, y in zip(q0, q1)], np.int32)
            print(f"m={m} n={n} kind={kind} queries={q0.size}", flush=True)
            assert np.array_equal(ctx.semilocal(r2c, m, n, kind, q0, q1), exp)
            assert np.array_equal(idx.query(m, n, kind, q0, q1), exp)
        idx.close()


def test_dominance_counts_brute_force(ctx):
    rng = np.random.default_rng(5)
    p = rng.permutation(3000).astype(np.int32)
    qr = rng.integers(0, 3001, 500).astype(np.int32)
    qc = rng.integers(0, 3001, 500).astype(np.int32)
    got = ctx.dominance(p, qr, qc)
    exp = np.array([(p[r:] < c).sum() for r, c in zip(qr, qc)])
    assert np.array_equal(got, exp)
    idx = lcs_amd.DominanceIndex(ctx, p)
    assert np.array_equal(idx.query(p.size, 0, lcs_amd.SEMI_DOMINANCE, qr, qc), exp)
    idx.close()


@pytest.mark.parametrize("size", [1, 2, 63, 64, 65, 127, 128, 4095, 4096, 4097, 70001])
def test_index_edge_sizes_with_unset_rows(ctx, size):
    rng = np.random.default_rng(size)
    p = rng.permutation(size).astype(np.int32)
    p[rng.random(size) < 0.1] = -1
    r0 = rng.integers(-3, size + 3, 3000).astype(np.int32)
    c0 = np.concatenate([rng.integers(-3, size + 3, 2998), [np.iinfo(np.int32).max, np.iinfo(np.int32).min]])
    c0 = c0.astype(np.int32)
    print(f"size={size}", flush=True)
    idx = lcs_amd.DominanceIndex(ctx, p)
    got = idx.query(size, 0, lcs_amd.SEMI_DOMINANCE, r0, c0)
    exp = oracle.semilocal(p, size, 0, lcs_amd.SEMI_DOMINANCE, r0, c0)
    assert np.array_equal(got, exp)
    idx.close()


def test_index_rejects_values_outside_rows(ctx):
    p = np.arange(100, dtype=np.int32)
    p[7] = 100
    with pytest.raises(lcs_amd.LcsError, match="EINVAL"):
        lcs_amd.DominanceIndex(ctx, p)
    p[7] = -2
    with pytest.raises(lcs_amd.LcsError, match="EINVAL"):
        lcs_amd.DominanceIndex(ctx, p)
    # the one-shot entry still answers (its direct count takes any int32),
    # also when the query set is large enough to prefer the index
    rng = np.random.default_rng(1)
    big = rng.permutation(1 << 20).astype(np.int32)
    big[5] = 1 << 21
    qr, qc = rng.integers(0, 1 << 20, 1500).astype(np.int32), rng.integers(0, 1 << 21, 1500).astype(np.int32)
    got = ctx.dominance(big, qr, qc)
    exp = oracle.semilocal(big, big.size, 0, lcs_amd.SEMI_DOMINANCE, qr[:200], qc[:200])
    assert np.array_equal(got[:200], exp)


def test_invalid_queries_are_rejected(ctx):
    a = np.array([1, 2, 3], np.int32)
    b = np.array([3, 1, 2, 2], np.int32)
    h, v, _ = ctx.comb(a, b)
    r2c = lcs_amd.from_strands(h, v)[0]
    idx = lcs_amd.DominanceIndex(ctx, r2c)
    for kind, q0, q1 in [(0, 2, 1), (0, 0, 5), (1, 0, 4), (1, -1, 2), (2, 4, 0), (3, 0, 5), (7, 0, 0)]:
        with pytest.raises(lcs_amd.LcsError, match="EINVAL"):
            ctx.semilocal(r2c, 3, 4, kind, [q0], [q1])
        with pytest.raises(lcs_amd.LcsError, match="EINVAL"):
            idx.query(3, 4, kind, [q0], [q1])
    with pytest.raises(lcs_amd.LcsError, match="EINVAL"):
        idx.query(3, 5, 0, [0], [1])   # m + n differs from the index size
    idx.close()


def test_large_pair_all_kinds_index_scan_and_oracle(ctx):
    """P of a 2^16 x 2^16 mutated pair: 20k queries per kind through the
    index, the one-shot entry (which builds its own index at this size) and
    the direct count; a sample against the oracle."""
    w = lcs_amd.workload_single("mutated2^16", seed=3)
    m, n = w.a.size, w.b.size
    h, v, llcs = ctx.comb(w.a, w.b)
    r2c = lcs_amd.from_strands(h, v)[0]
    idx = lcs_amd.DominanceIndex(ctx, r2c)
    rng = np.random.default_rng(77)
    for kind in KINDS + (lcs_amd.SEMI_DOMINANCE,):
        q0, q1 = random_queries(rng, kind, m, n, 20000)
        print(f"kind={kind}", flush=True)
        got = idx.query(m, n, kind, q0, q1)
        assert np.array_equal(ctx.semilocal(r2c, m, n, kind, q0, q1), got)   # one-shot: index path
        assert np.array_equal(ctx.semilocal(r2c, m, n, kind, q0[:200], q1[:200]), got[:200])   # direct count
        assert np.array_equal(oracle.semilocal(r2c, m, n, kind, q0[:300], q1[:300]), got[:300])
    full = idx.query(m, n, lcs_amd.SEMI_STRING_SUBSTRING, [0], [n])
    assert full[0] == llcs
    assert idx.query(m, n, lcs_amd.SEMI_SUBSTRING_STRING, [0], [m])[0] == llcs
    assert idx.query(m, n, lcs_amd.SEMI_PREFIX_SUFFIX, [m], [0])[0] == llcs
    assert idx.query(m, n, lcs_amd.SEMI_SUFFIX_PREFIX, [0], [n])[0] == llcs
    assert idx.query(m, n, lcs_amd.SEMI_STRING_SUBSTRING, [100], [700])[0] == oracle.llcs_dp(w.a, w.b[100:700])
    idx.close()


def test_device_index_and_device_queries(ctx):
    """Build from a device permutation, query from device buffers; a query
    outside its quadrant yields -1 on this path."""
    w = lcs_amd.workload_single("uniform4096", seed=2)
    m, n = w.a.size, w.b.size
    h, v, _ = ctx.comb(w.a, w.b)
    r2c = lcs_amd.from_strands(h, v)[0]
    d_p = H.DeviceArray.from_numpy(r2c)
    idx = lcs_amd.DominanceIndex(ctx, size=r2c.size, device_ptr=d_p.ptr.value)
    d_p.free()   # the index keeps no reference to its input
    rng = np.random.default_rng(9)
    q0, q1 = random_queries(rng, lcs_amd.SEMI_PREFIX_SUFFIX, m, n, 5000)
    q0[17], q1[17] = m + 1, 0
    d0, d1, dout = H.DeviceArray.from_numpy(q0), H.DeviceArray.from_numpy(q1), H.DeviceArray(q0.size)
    idx.query_device(m, n, lcs_amd.SEMI_PREFIX_SUFFIX, d0.ptr.value, d1.ptr.value, q0.size, dout.ptr.value)
    H.synchronize()
    got = dout.numpy()
    exp = oracle.semilocal(r2c, m, n, lcs_amd.SEMI_PREFIX_SUFFIX, q0, q1)
    assert got[17] == -1 and exp[17] == -1
    assert np.array_equal(got, exp)
    assert idx.nbytes >= 13 * (r2c.size // 64 + 1) * 16
    idx.close()


def test_windows_of_a_large_pair(ctx):
    w = lcs_amd.workload_single("mutated2^14", seed=9)
    h, v, llcs = ctx.comb(w.a, w.b)
    r2c = lcs_amd.from_strands(h, v)[0]
    n = w.b.size
    got = ctx.string_substring(r2c, w.a.size, n, [0, 100, 5000, 0], [n, 700, 9000, 0])
    assert got[0] == llcs and got[3] == 0
    assert got[1] == oracle.llcs_dp(w.a, w.b[100:700])
