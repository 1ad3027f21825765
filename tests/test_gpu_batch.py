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
    pairs = lcs_amd.workload_batch(40, seed=21, lo=1, hi=5000)
    scheds = [lcs_amd.schedule(tile_m=[1, 2, 4, 8][i % 4], variant=1 + (i % 2)) for i in range(len(pairs))]
    hs, vs = ctx.batch_comb([p.a for p in pairs], [p.b for p in pairs], scheds=scheds)
    for p, h, v in zip(pairs, hs, vs):
        ho, vo = oracle.comb(p.a, p.b)
        assert np.array_equal(h, ho) and np.array_equal(v, vo)
