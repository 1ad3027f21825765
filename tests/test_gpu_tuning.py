"""The learned schedule cost model: what it gains over the method without it.

The model is trained on the GPU (lcs_mlp_train) on the autotuner's timing
table recorded on this device (committed under lcs_amd/data).  Held out by
k-fold, its choice is scored by the exhaustive autotune's own measurements
(regret = chosen cost / best cost) and compared with the built-in heuristic.
It must pick schedules at least as good as the heuristic on average, and it
decides in well under a millisecond where the exhaustive autotune spends
seconds per shape.  Schedules never change results: a batch run with the
model's choices is bit-identical to one with the heuristic's."""
import os
import time

import numpy as np
import pytest

import lcs_amd
import oracle
from lcs_amd import tuning

pytestmark = pytest.mark.gpu


def _table(mode):
    if not os.path.exists(tuning.TABLES[mode]):
        pytest.skip(f"no autotune table for mode {mode} on this device")
    return tuning.load_table(mode)


@pytest.mark.parametrize("mode", [1, 0])
def test_model_beats_heuristic_on_held_out_shapes(ctx, mode):
    table = _table(mode)
    rep = tuning.evaluate(ctx, table, folds=4, steps=3000)
    print(rep)
    assert rep["model_regret_mean"] <= rep["heuristic_regret_mean"] + 1e-9
    assert rep["model_regret_mean"] < 1.25


def test_model_decides_far_faster_than_exhaustive_tuning(ctx):
    table = _table(1)
    mlp = tuning.train(ctx, table, steps=2000)
    X = tuning.features([r["m"] for r in table["rows"]], [r["n"] for r in table["rows"]],
                        [r["alphabet"] for r in table["rows"]], [r["rate"] for r in table["rows"]])
    mlp.choose(X)   # warm
    t0 = time.perf_counter()
    for _ in range(10):
        mlp.choose(X)
    per_shape = (time.perf_counter() - t0) / 10 / len(X)
    tune_s = np.mean([r["tune_s"] for r in table["rows"]])
    assert per_shape * 100 < tune_s
    mlp.close()


def test_learned_schedules_do_not_change_results(ctx):
    table = _table(1)
    mlp = tuning.train(ctx, table, steps=1000)
    pairs = lcs_amd.workload_batch(64, seed=31, lo=256, hi=4096)
    scheds, _ = tuning.choose_schedules(mlp, [p.a.size for p in pairs], [p.b.size for p in pairs],
                                        [p.alphabet for p in pairs], [p.rate for p in pairs],
                                        tuning.table_candidates(table))
    h1, v1 = ctx.batch_comb([p.a for p in pairs], [p.b for p in pairs], scheds=scheds)
    h2, v2 = ctx.batch_comb([p.a for p in pairs], [p.b for p in pairs])
    assert all(np.array_equal(x, y) for x, y in zip(h1, h2))
    assert all(np.array_equal(x, y) for x, y in zip(v1, v2))
    mlp.close()


def test_autotune_measure_small_shape_and_its_choice_is_exact(ctx):
    """The autotuner itself on a small shape: every candidate's cost is
    finite and positive in both modes, and the cheapest schedule combs the
    pair bit-for-bit like the oracle."""
    m, n = 1500, 3000
    for mode in (0, 1):
        cost = ctx.autotune(m, n, 4, 0.1, tuning.CANDIDATES, mode=mode, reps=1, seed=5)
        assert cost.shape == (len(tuning.CANDIDATES),)
        assert np.all(np.isfinite(cost)) and np.all(cost > 0), (mode, cost)
        best = tuning.CANDIDATES[int(np.argmin(cost))]
        w = lcs_amd.workload_batch(1, seed=40 + mode, lo=m, hi=m)[0]
        h, v, _ = ctx.comb(w.a, w.b, best)
        ho, vo = oracle.comb(w.a, w.b, order="simd")
        assert np.array_equal(h, ho) and np.array_equal(v, vo), (mode, best)


def _same_schedule(s, t):
    return (s.tile_m, s.waves_per_simd, s.variant, s.strips) == (t.tile_m, t.waves_per_simd, t.variant, t.strips)


def test_selection_order_table_model_autotune_default():
    """lcs_ctx_choose_schedule walks table -> model -> autotune on a miss ->
    default, and lcs_comb / lcs_batch_comb with no schedule follow it with
    results equal to the oracle's."""
    with lcs_amd.Context(0) as c:
        for mode in (lcs_amd.TUNE_SINGLE, lcs_amd.TUNE_BATCH):
            s, src = c.choose_schedule(mode, 3000, 5000)
            assert src == "default" and s.tile_m == 0
        # 1. a table row answers for its shape (and shapes within 1/8 octave)
        row = lcs_amd.schedule(tile_m=2, waves_per_simd=1, variant=2, strips=1)
        c.add_tuned(lcs_amd.TUNE_SINGLE, 3000, 5000, row)
        s, src = c.choose_schedule(lcs_amd.TUNE_SINGLE, 3100, 5200)
        assert src == "table" and _same_schedule(s, row)
        s, src = c.choose_schedule(lcs_amd.TUNE_SINGLE, 6000, 5000)   # an octave away: a miss
        assert src == "default"
        # 2. a model answers what the table leaves open; the table still wins
        table = tuning.load_table(1) if os.path.exists(tuning.TABLES[1]) else None
        if table is None:
            pytest.skip("no autotune table to train the model on")
        mlp = tuning.train(c, table, steps=500)
        cands = tuning.table_candidates(table)
        c.set_cost_model(lcs_amd.TUNE_SINGLE, mlp, cands)
        s, src = c.choose_schedule(lcs_amd.TUNE_SINGLE, 6000, 5000, alphabet=4, rate=0.1)
        expect = cands[int(mlp.choose(tuning.features([6000], [5000], [4], [0.1]))[0])]
        assert src == "model" and _same_schedule(s, expect)
        s, src = c.choose_schedule(lcs_amd.TUNE_SINGLE, 3000, 5000)
        assert src == "table" and _same_schedule(s, row)
        # every source gives the oracle's strands
        for shape_seed, (mm, nn) in enumerate([(3000, 5000), (6000, 5000)]):
            w = lcs_amd.workload_batch(1, seed=60 + shape_seed, lo=mm, hi=mm)[0]
            b = w.b[:nn] if w.b.size >= nn else np.concatenate([w.b, w.a[: nn - w.b.size]])
            h, v, _ = c.comb(w.a, b)
            ho, vo = oracle.comb(w.a, b, order="simd")
            assert np.array_equal(h, ho) and np.array_equal(v, vo)
        # destroying the model detaches it
        mlp.close()
        s, src = c.choose_schedule(lcs_amd.TUNE_SINGLE, 6000, 5000)
        assert src == "default"
        # 3. the exhaustive autotune on a miss; its result joins the table
        cands = tuning.CANDIDATES[:3]
        c.set_autotune_on_miss(lcs_amd.TUNE_SINGLE, cands, reps=1)
        s, src = c.choose_schedule(lcs_amd.TUNE_SINGLE, 700, 900)
        assert src == "autotune" and any(_same_schedule(s, x) for x in cands)
        s2, src2 = c.choose_schedule(lcs_amd.TUNE_SINGLE, 700, 900)
        assert src2 == "table" and _same_schedule(s, s2)
        c.set_autotune_on_miss(lcs_amd.TUNE_SINGLE, [], reps=1)
        # batch mode: table rows per pair, default for the rest, same bits
        c.add_tuned(lcs_amd.TUNE_BATCH, 1000, 1000, lcs_amd.schedule(tile_m=4, waves_per_simd=1, variant=1))
        pairs = lcs_amd.workload_batch(24, seed=71, lo=800, hi=1300)
        hs, vs = c.batch_comb([p.a for p in pairs], [p.b for p in pairs])
        for p, h, v in zip(pairs, hs, vs):
            ho, vo = oracle.comb(p.a, p.b, order="simd")
            assert np.array_equal(h, ho) and np.array_equal(v, vo)


def test_loaded_tables_pick_their_measured_best(ctx):
    """The committed tables loaded into a context: every recorded shape
    gets its measured best schedule from step 1 (the table)."""
    with lcs_amd.Context(0) as c:
        tuning.load_into(c)
        for mode in (0, 1):
            if not os.path.exists(tuning.TABLES[mode]):
                continue
            for m, n, best in tuning.best_rows(tuning.load_table(mode))[:8]:
                s, src = c.choose_schedule(mode, m, n)
                assert src == "table" and _same_schedule(s, best)
