"""The batched path combs every pair exactly as the single-pair path and
the oracle do, with and without the hipGraph capture."""
import numpy as np
import pytest

import lcs_amd
import oracle

pytestmark = pytest.mark.gpu


def test_batch_uneven_pairs_match_oracle(ctx):
    pairs = lcs_amd.workload_batch(60, seed=11, lo=1, hi=3000)
    a_list = [p.a for p in pairs]
    b_list = [p.b for p in pairs]
    for use_graph in (True, False):
        hs, vs = ctx.batch_comb(a_list, b_list, use_graph=use_graph)
        for p, h, v in zip(pairs, hs, vs):
            ho, vo = oracle.comb(p.a, p.b)
            assert np.array_equal(h, ho) and np.array_equal(v, vo)


def test_batch_with_explicit_schedules_and_empty_pairs(ctx):
    rng = np.random.default_rng(3)
    a_list = [rng.integers(0, 4, m).astype(np.int32) for m in (0, 5, 700, 0, 1300, 64, 2000)]
    b_list = [rng.integers(0, 4, n).astype(np.int32) for n in (3, 0, 900, 0, 64, 1300, 17)]
    tms = [1, 2, 4, 8, 16, 32, 4]
    scheds = [lcs_amd.schedule(tile_m=t) for t in tms]
    hs, vs = ctx.batch_comb(a_list, b_list, scheds=scheds)
    for a, b, h, v in zip(a_list, b_list, hs, vs):
        ho, vo = oracle.comb(a, b)
        assert np.array_equal(h, ho) and np.array_equal(v, vo)


def test_batch_equals_single_calls(ctx):
    pairs = lcs_amd.workload_batch(12, seed=5)
    hs, vs = ctx.batch_comb([p.a for p in pairs], [p.b for p in pairs])
    for p, h, v in zip(pairs, hs, vs):
        h1, v1, _ = ctx.comb(p.a, p.b)
        assert np.array_equal(h, h1) and np.array_equal(v, v1)


def test_batch_mixed_variants(ctx):
    # register chain, LDS groups, and the byte-chain variant (in a batch the
    # register chain with 32-bit symbols)
    pairs = lcs_amd.workload_batch(40, seed=21, lo=1, hi=5000)
    variants = [lcs_amd.VARIANT_CHAIN, lcs_amd.VARIANT_CHAIN_LDS, lcs_amd.VARIANT_CHAIN_BYTES]
    scheds = [lcs_amd.schedule(tile_m=[1, 2, 4, 8][i % 4], variant=variants[i % 3]) for i in range(len(pairs))]
    hs, vs = ctx.batch_comb([p.a for p in pairs], [p.b for p in pairs], scheds=scheds)
    for p, h, v in zip(pairs, hs, vs):
        ho, vo = oracle.comb(p.a, p.b)
        assert np.array_equal(h, ho) and np.array_equal(v, vo)


def test_batch_graph_reused_across_calls(ctx):
    """The instantiated graph of one batch is updated in place for the next
    batch with the same classes and rebuilt when the classes change; every
    call, with eager single-pair combs in between, equals the oracle."""
    sets = [lcs_amd.workload_batch(30, seed=s, lo=300, hi=4000) for s in (41, 42)]
    odd = lcs_amd.workload_batch(10, seed=43, lo=20, hi=60)            # other tile classes
    ref = {}
    for seq, pairs in enumerate([sets[0], sets[1], sets[0], odd, sets[1], sets[0]]):
        hs, vs = ctx.batch_comb([p.a for p in pairs], [p.b for p in pairs], use_graph=True)
        for i, (p, h, v) in enumerate(zip(pairs, hs, vs)):
            key = (id(pairs), i)
            if key not in ref:
                ref[key] = oracle.comb(p.a, p.b, order="simd")
            ho, vo = ref[key]
            assert np.array_equal(h, ho) and np.array_equal(v, vo), (seq, i)
        h1, v1, _ = ctx.comb(pairs[0].a, pairs[0].b)                 # eager work between replays
        assert np.array_equal(h1, ref[(id(pairs), 0)][0])


PACKED16 = 3  # LCS_VARIANT_PACKED16


def test_packed16_batch_matches_oracle(ctx):
    """Pairs combed two per wave with 16-bit strands equal the oracle: uneven
    shapes paired within each tile height (odd ones out on the 32-bit chain),
    symbols at the top of the 16-bit range, and pairs the packed chain must
    refuse (m + n > 65535, a symbol outside [0, 65533], empty strings)."""
    rng = np.random.default_rng(16)
    pairs = lcs_amd.workload_batch(61, seed=61, lo=1, hi=6000)
    a_list, b_list = [p.a for p in pairs], [p.b for p in pairs]
    # symbols near the top of the range, and an alphabet of 65534 symbols
    a_list.append(rng.integers(65500, 65534, 3000).astype(np.int32))
    b_list.append(rng.integers(65500, 65534, 2500).astype(np.int32))
    a_list.append(rng.integers(0, 65534, 5000).astype(np.int32))
    b_list.append(np.concatenate([a_list[-1][1000:4000], rng.integers(0, 65534, 900)]).astype(np.int32))
    # refused by the packed chain: too many strands, symbols out of range, empty
    a_list.append(rng.integers(0, 4, 40000).astype(np.int32))
    b_list.append(rng.integers(0, 4, 30000).astype(np.int32))
    a_list.append(np.array([70000, 1, 2, 3] * 300, np.int32))
    b_list.append(np.array([1, 2, 3, 70000] * 250, np.int32))
    a_list.append(np.array([-5, 1, 2] * 100, np.int32))
    b_list.append(np.array([1, -5] * 100, np.int32))
    a_list.append(np.zeros(0, np.int32))
    b_list.append(rng.integers(0, 4, 77).astype(np.int32))
    tms = [1, 2, 4, 8, 16, 32]
    scheds = [lcs_amd.schedule(tile_m=tms[i % 6] if i % 7 else 0, variant=PACKED16) for i in range(len(a_list))]
    hs, vs = ctx.batch_comb(a_list, b_list, scheds=scheds)
    for i, (a, b, h, v) in enumerate(zip(a_list, b_list, hs, vs)):
        ho, vo = oracle.comb(a, b, order="simd")
        assert np.array_equal(h, ho) and np.array_equal(v, vo), (i, len(a), len(b))


def test_packed16_onehot_and_xor_cells_match_oracle(ctx):
    """Couples whose symbols all lie in [0, 15] run the one-hot packed cells
    (1 << symbol per half), the others the xor cells: alphabets of exactly 16
    (symbol 15 present) and 17 (symbol 16 present, so the pair is refused by
    the one-hot cells), a batch mixing both at every tile height so that one
    couple joins a one-hot pair with a wide one, uneven shapes (padding rows
    and columns carry 0 in the one-hot encoding), and long runs of a single
    symbol; every pair equals the oracle."""
    rng = np.random.default_rng(1616)
    a_list, b_list, tms = [], [], []
    for tm in (1, 2, 4, 8, 16):
        for k, sigma in enumerate((16, 17, 4, 16, 17, 2, 16)):
            m = int(rng.integers(40, 2500)); n = int(rng.integers(40, 2500))
            a = rng.integers(0, sigma, m).astype(np.int32)
            b = rng.integers(0, sigma, n).astype(np.int32)
            a[0] = sigma - 1                     # the top symbol is present
            if k == 5:                           # long single-symbol runs
                a[: m // 2] = 1
                b[n // 3:] = 1
            a_list.append(a); b_list.append(b); tms.append(tm)
    scheds = [lcs_amd.schedule(tile_m=t, variant=PACKED16) for t in tms]
    hs, vs = ctx.batch_comb(a_list, b_list, scheds=scheds)
    for i, (a, b, h, v) in enumerate(zip(a_list, b_list, hs, vs)):
        ho, vo = oracle.comb(a, b, order="simd")
        assert np.array_equal(h, ho) and np.array_equal(v, vo), (i, tms[i], len(a), len(b), int(a.max()))


def test_packed16_large_batch_equals_32bit(ctx):
    """A config-4-like batch (lengths log-uniform in [256, 16384]) through the
    packed chain, bit for bit against the 32-bit register chain, and the
    library's own default (no schedules: the packed chain for every pair it
    can take) equal to both."""
    pairs = lcs_amd.workload_batch(2000, seed=77)
    a_list, b_list = [p.a for p in pairs], [p.b for p in pairs]
    h32, v32 = ctx.batch_comb(a_list, b_list, scheds=[lcs_amd.schedule(variant=lcs_amd.VARIANT_CHAIN)] * len(pairs))
    h16, v16 = ctx.batch_comb(a_list, b_list, scheds=[lcs_amd.schedule(variant=PACKED16)] * len(pairs))
    bad = [i for i in range(len(pairs)) if not (np.array_equal(h16[i], h32[i]) and np.array_equal(v16[i], v32[i]))]
    assert not bad, bad[:5]
    hd, vd = ctx.batch_comb(a_list, b_list)
    bad = [i for i in range(len(pairs)) if not (np.array_equal(hd[i], h32[i]) and np.array_equal(vd[i], v32[i]))]
    assert not bad, bad[:5]


def test_batch_strips_merged_in_the_graph(ctx):
    """Pairs with strips > 1 in a batch are combed as column strips in the
    batch's chain launches and merged by the steady ant in the same captured
    graph (k = m, several products per level, big and small cores, odd strip
    counts); every pair equals the oracle, with and without the graph."""
    rng = np.random.default_rng(90)
    pairs = lcs_amd.workload_batch(40, seed=91, lo=200, hi=8000)
    a_list, b_list = [p.a for p in pairs], [p.b for p in pairs]
    strips = [0] * len(pairs)
    for i in range(0, len(pairs), 3):
        strips[i] = [2, 3, 4, 5, 8][i % 5]
    # a tall pair whose merges run on the whole-GPU steady ant (k = 40000)
    a_list.append(rng.integers(0, 4, 40000).astype(np.int32))
    b_list.append(rng.integers(0, 4, 90000).astype(np.int32))
    strips.append(3)
    a_list.append(rng.integers(0, 4, 3).astype(np.int32))      # more strips than columns
    b_list.append(rng.integers(0, 4, 2).astype(np.int32))
    strips.append(8)
    scheds = [lcs_amd.schedule(strips=s) for s in strips]
    ref = [oracle.comb(a, b, order="simd") for a, b in zip(a_list, b_list)]
    for use_graph in (True, False, True):
        hs, vs = ctx.batch_comb(a_list, b_list, scheds=scheds, use_graph=use_graph)
        for i, (h, v, (ho, vo)) in enumerate(zip(hs, vs, ref)):
            assert np.array_equal(h, ho) and np.array_equal(v, vo), (use_graph, i, strips[i])
