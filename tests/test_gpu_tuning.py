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
                                        [p.alphabet for p in pairs], [p.rate for p in pairs])
    h1, v1 = ctx.batch_comb([p.a for p in pairs], [p.b for p in pairs], scheds=scheds)
    h2, v2 = ctx.batch_comb([p.a for p in pairs], [p.b for p in pairs])
    assert all(np.array_equal(x, y) for x, y in zip(h1, h2))
    assert all(np.array_equal(x, y) for x, y in zip(v1, v2))
    mlp.close()
