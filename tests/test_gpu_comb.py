"""Parity of the HIP band-chain comb with the CPU oracle (bit-exact)."""
import numpy as np
import pytest

import lcs_amd
import oracle

pytestmark = pytest.mark.gpu
BYTES = lcs_amd.VARIANT_CHAIN_BYTES   # the register chain, byte symbols staged at any size when b allows them


def _check(ctx, a, b, sched=None):
    h, v, llcs = ctx.comb(a, b, sched)
    ho, vo = oracle.comb(a, b)
    assert np.array_equal(h, ho), f"h mismatch m={len(a)} n={len(b)} sched={sched}"
    assert np.array_equal(v, vo), f"v mismatch m={len(a)} n={len(b)} sched={sched}"
    assert llcs == oracle.llcs_from_strands(vo, len(a))
    return h, v


def test_tiny_known_answer(ctx):
    h, v, llcs = ctx.comb([1, 2], [1])
    assert list(h) == [0, 2] and list(v) == [1] and llcs == 1


@pytest.mark.parametrize("m,n", [(0, 0), (0, 5), (7, 0), (1, 1), (1, 100), (100, 1), (63, 64), (64, 63),
                                 (65, 129), (300, 17), (17, 300)])
def test_edge_shapes(ctx, m, n):
    rng = np.random.default_rng(m * 1000 + n)
    a = rng.integers(0, 3, m).astype(np.int32)
    b = rng.integers(0, 3, n).astype(np.int32)
    for tm in (1, 4, 16):
        _check(ctx, a, b, lcs_amd.schedule(tile_m=tm))
        _check(ctx, a, b, lcs_amd.schedule(tile_m=tm, variant=BYTES))


def test_interior_chunk_bounds(ctx):
    """Widths around the band loop's phase boundaries: a chunk of the
    register chain is 32 steps, the interior (unmasked) chunks start at step
    96 and stop two chunks before n, and run two per iteration, so every
    n mod 64 and the widths with zero, one, two and three interior chunks
    are covered, across several bands, for the register chain and for the
    LDS band groups (same chunk length, their own loop)."""
    rng = np.random.default_rng(404)
    for n in list(range(76, 100)) + [111, 112, 113, 127, 128, 129, 143, 144, 145, 159, 160, 161, 175, 176, 177,
                                     191, 192, 193, 207, 208, 209, 223, 224, 225, 255, 256, 257, 287, 288, 289,
                                     319, 320, 321] + list(range(960, 1026)):
        a = rng.integers(0, 3, 2500).astype(np.int32)
        b = rng.integers(0, 3, n).astype(np.int32)
        for tm in (4, 16):
            _check(ctx, a, b, lcs_amd.schedule(tile_m=tm))
            _check(ctx, a, b, lcs_amd.schedule(tile_m=tm, variant=BYTES))
        _check(ctx, a, b, lcs_amd.schedule(tile_m=4, variant=LDS))


@pytest.mark.parametrize("tm", [1, 2, 4, 8, 16, 32])
def test_tile_heights(ctx, tm):
    rng = np.random.default_rng(tm)
    for m, n in [(1000, 1500), (4096, 777), (5000, 64 * 32 + 5)]:
        a = rng.integers(0, 4, m).astype(np.int32)
        b = rng.integers(0, 4, n).astype(np.int32)
        _check(ctx, a, b, lcs_amd.schedule(tile_m=tm))
        _check(ctx, a, b, lcs_amd.schedule(tile_m=tm, variant=BYTES))


def test_uniform4096(ctx):
    w = lcs_amd.workload_single("uniform4096")
    for tm in (0, 1, 2, 4):
        _check(ctx, w.a, w.b, lcs_amd.schedule(tile_m=tm))


@pytest.mark.parametrize("m,n", [(7330, 9944), (9771, 11302)])
def test_reference_script_sizes(ctx, m, n):
    # non-aligned sizes of the reference's run script (test_default.txt)
    a = lcs_amd.gen_binary(888, m)
    b = lcs_amd.gen_binary(999, n)
    _check(ctx, a, b)
    _check(ctx, a, b, lcs_amd.schedule(tile_m=8, waves_per_simd=2))


def test_large_alphabet_and_negative_symbols(ctx):
    rng = np.random.default_rng(5)
    a = rng.integers(-2**31, 2**31 - 1, 3000, dtype=np.int64).astype(np.int32)
    b = rng.integers(-2**31, 2**31 - 1, 2000, dtype=np.int64).astype(np.int32)
    b[::7] = a[: len(b[::7])]
    _check(ctx, a, b)


def test_byte_symbols_and_their_range_check(ctx):
    """The register chain reads byte copies of b when every symbol of b lies
    in [0, 255] (a device-side check per call) and b itself otherwise (the
    byte-chain variant stages the copies at any size).  Row symbols
    congruent to a column symbol mod 256 must not match it, a single symbol
    outside the byte range anywhere in b must switch the call to the 32-bit
    symbols, and the next call with byte symbols must switch back; strips
    read the copies at their own column offsets."""
    rng = np.random.default_rng(256)
    m, n = 3000, 5003
    b = rng.integers(0, 256, n).astype(np.int32)
    a = rng.integers(0, 256, m).astype(np.int32)
    a[::3] += 256                                   # never equal to a byte symbol
    for tm in (1, 2, 4, 8, 16, 32):
        _check(ctx, a, b, lcs_amd.schedule(tile_m=tm, variant=BYTES))
    for last in (256, -1, 1 << 20):
        bb = b.copy()
        bb[-1] = last
        aa = a.copy()
        aa[::5] = last                              # matches only in 32 bits
        _check(ctx, aa, bb, lcs_amd.schedule(tile_m=16, variant=BYTES))
        _check(ctx, a, b, lcs_amd.schedule(tile_m=16, variant=BYTES))
    bb = b.copy()
    bb[0] = 300
    _check(ctx, a, bb, lcs_amd.schedule(tile_m=4, strips=3, variant=BYTES))
    for strips in (2, 3, 7):                        # strip offsets of every residue mod 4
        _check(ctx, a, b, lcs_amd.schedule(tile_m=4, strips=strips, variant=BYTES))
        _check(ctx, a, b, lcs_amd.schedule(tile_m=16, strips=strips, variant=BYTES))


def test_all_match_and_no_match(ctx):
    a = np.zeros(2000, np.int32)
    h, v, llcs = ctx.comb(a, np.zeros(1500, np.int32))
    assert llcs == 1500
    h, v, llcs = ctx.comb(a, np.ones(1500, np.int32))
    assert llcs == 0 and np.array_equal(h, np.arange(2000)) and np.array_equal(v, 2000 + np.arange(1500))


def test_mutated_pair_and_hash(ctx):
    w = lcs_amd.workload_single("mutated2^15")
    h, v = _check(ctx, w.a, w.b)
    r2c, c2r = lcs_amd.from_strands(h, v)
    assert lcs_amd.perm_hash(c2r) == oracle.perm_hash(oracle.from_strands(h, v)[1])


def test_fixup_quadrants(ctx):
    rng = np.random.default_rng(11)
    m, n = 700, 900
    a = rng.integers(0, 4, m).astype(np.int32)
    b = rng.integers(0, 4, n).astype(np.int32)
    ap = a[::-1].copy()
    h = np.arange(m, dtype=np.int32)
    v = (m + np.arange(n)).astype(np.int32)
    # reference order: rows with high i first (they feed v down), then the rest
    i_split, j_split = 300, 450
    ctx.comb_fixup(ap, b, h, v, i_split, m - i_split, 0, j_split)
    ctx.comb_fixup(ap, b, h, v, 0, i_split, 0, j_split)
    ctx.comb_fixup(ap, b, h, v, 0, m, j_split, n - j_split)
    ho, vo = oracle.comb(a, b)
    assert np.array_equal(h, ho) and np.array_equal(v, vo)


def test_repeatable(ctx):
    w = lcs_amd.workload_single("mutated2^14")
    r1 = ctx.comb(w.a, w.b)
    r2 = ctx.comb(w.a, w.b, lcs_amd.schedule(tile_m=4, waves_per_simd=2))
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])


# 80 strips: the first merge level has 40 products, more than one slicing
# launch takes (32), and later levels carry an odd permutation over
@pytest.mark.parametrize("strips", [2, 3, 8, 80])
def test_column_strips_merged_by_steady_ant(ctx, strips):
    rng = np.random.default_rng(strips)
    a = rng.integers(0, 4, 3000).astype(np.int32)
    b = rng.integers(0, 4, 40000).astype(np.int32)
    _check(ctx, a, b, lcs_amd.schedule(tile_m=4, strips=strips))


def test_thin_pair_default_schedule(ctx):
    # a thin, wide pair like the Zipf workload (reduced): the default
    # schedule cuts it into strips and merges them on the GPU
    a = lcs_amd.gen_zipf(5, 1 << 12, 26, 1.1)
    b = lcs_amd.gen_planted(6, a, 1 << 18, 26, 1.1, 16, 0.05)
    h, v, _ = ctx.comb(a, b)
    ho, vo = oracle.comb(a, b, order="mt", threads=8)
    assert np.array_equal(h, ho) and np.array_equal(v, vo)


def test_thin_pair_eight_row_strips(ctx):
    # m = 2^14, n = 2^19: the default schedule combs 32 strips of 32 bands of
    # 8 rows per lane and merges them with five steady-ant levels; the result
    # must equal the pair combed as one strip (itself pinned to the oracle by
    # the tests above)
    a = lcs_amd.gen_zipf(11, 1 << 14, 26, 1.1)
    b = lcs_amd.gen_planted(12, a, 1 << 19, 26, 1.1, 8, 0.05)
    h, v, llcs = ctx.comb(a, b)
    h1, v1, llcs1 = ctx.comb(a, b, lcs_amd.schedule(strips=1))
    assert np.array_equal(h, h1) and np.array_equal(v, v1) and llcs == llcs1


def test_thin_pair_at_the_largest_strip_height(ctx):
    # m = 32768, the largest height the default schedule cuts into strips;
    # one row more and the default combs one strip, while explicit strips
    # merge on the whole-GPU steady ant (k = m = 32769) to the same strands
    a = lcs_amd.gen_zipf(13, 1 << 15, 4, 1.1)
    b = lcs_amd.gen_planted(14, a, 1 << 18, 4, 1.1, 4, 0.05)
    h, v, llcs = ctx.comb(a, b)
    h1, v1, llcs1 = ctx.comb(a, b, lcs_amd.schedule(strips=1))
    assert np.array_equal(h, h1) and np.array_equal(v, v1) and llcs == llcs1
    a2 = np.append(a, 1).astype(np.int32)
    h2, v2, _ = ctx.comb(a2, b, lcs_amd.schedule(strips=2))
    h3, v3, _ = ctx.comb(a2, b, lcs_amd.schedule(strips=1))
    assert np.array_equal(h2, h3) and np.array_equal(v2, v3)


LDS = 2  # LCS_VARIANT_CHAIN_LDS


@pytest.mark.parametrize("m,n", [(1, 1), (1, 100), (100, 1), (63, 64), (65, 129), (300, 17), (17, 300),
                                 (64 * 4 * 4, 500), (64 * 4 * 4 + 1, 500), (64 * 5, 1000)])
def test_lds_groups_edge_shapes(ctx, m, n):
    # groups of 4 bands: full groups, partial last groups (1..3 bands) and
    # a partial last band
    rng = np.random.default_rng(m * 7 + n)
    a = rng.integers(0, 3, m).astype(np.int32)
    b = rng.integers(0, 3, n).astype(np.int32)
    for tm in (1, 4):
        _check(ctx, a, b, lcs_amd.schedule(tile_m=tm, variant=LDS))


@pytest.mark.parametrize("tm", [1, 2, 4, 8, 16, 32])
def test_lds_groups_tile_heights(ctx, tm):
    rng = np.random.default_rng(100 + tm)
    for m, n in [(1000, 1500), (4096 * 3 + 11, 777), (5000, 64 * 32 + 5)]:
        a = rng.integers(0, 4, m).astype(np.int32)
        b = rng.integers(0, 4, n).astype(np.int32)
        _check(ctx, a, b, lcs_amd.schedule(tile_m=tm, variant=LDS))


def test_lds_groups_reference_sizes_and_strips(ctx):
    a = lcs_amd.gen_binary(888, 7330)
    b = lcs_amd.gen_binary(999, 9944)
    _check(ctx, a, b, lcs_amd.schedule(variant=LDS))
    _check(ctx, a, b, lcs_amd.schedule(tile_m=2, waves_per_simd=2, variant=LDS))
    rng = np.random.default_rng(9)
    a = rng.integers(0, 4, 3000).astype(np.int32)
    b = rng.integers(0, 4, 40000).astype(np.int32)
    _check(ctx, a, b, lcs_amd.schedule(tile_m=4, strips=4, variant=LDS))
    w = lcs_amd.workload_single("mutated2^14")
    r1 = ctx.comb(w.a, w.b)
    r2 = ctx.comb(w.a, w.b, lcs_amd.schedule(tile_m=4, variant=LDS))
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])


def test_unknown_variant_is_rejected(ctx):
    with pytest.raises(lcs_amd.LcsError):
        ctx.comb(np.zeros(10, np.int32), np.zeros(10, np.int32), lcs_amd.schedule(variant=7))


@pytest.mark.parametrize("m,n,strips", [(0, 0, 0), (0, 5, 0), (7, 0, 0), (1, 1, 0), (65, 129, 0), (7330, 9944, 0),
                                        (3000, 40000, 3), (3000, 40000, 80)])
def test_permutation_built_on_the_gpu(ctx, m, n, strips):
    """lcs_comb_permutation: the comb's strands turned into the permutation
    on the device must equal PermutationMatrix::FromStrands of the strands
    (the host routine, itself pinned to the oracle)."""
    rng = np.random.default_rng(m * 31 + n + strips)
    a = rng.integers(0, 4, m).astype(np.int32)
    b = rng.integers(0, 4, n).astype(np.int32)
    sched = lcs_amd.schedule(tile_m=4, strips=strips) if strips else None
    h, v, llcs = ctx.comb(a, b, sched)
    r2c, c2r, llcs2 = ctx.comb_permutation(a, b, sched)
    r2c_o, c2r_o = lcs_amd.from_strands(h, v)
    assert np.array_equal(r2c, r2c_o) and np.array_equal(c2r, c2r_o) and llcs2 == llcs
    if m + n:
        assert np.array_equal(np.sort(r2c), np.arange(m + n))


def test_permutation_built_on_the_gpu_thin_pair_hash(ctx):
    a = lcs_amd.gen_zipf(5, 1 << 12, 26, 1.1)
    b = lcs_amd.gen_planted(6, a, 1 << 18, 26, 1.1, 16, 0.05)
    r2c, c2r, llcs = ctx.comb_permutation(a, b)
    ho, vo = oracle.comb(a, b, order="mt", threads=8)
    assert lcs_amd.perm_hash(c2r) == oracle.perm_hash(oracle.from_strands(ho, vo)[1])
    assert llcs == oracle.llcs_from_strands(vo, len(a))
