"""Host-side logic of the tuning flow and the bench contract (no GPU)."""
import os
import subprocess
import sys

import numpy as np

from lcs_amd import tuning

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_candidates_cover_every_tile_and_variant():
    seen = {(c.tile_m, c.variant) for c in tuning.CANDIDATES}
    assert seen == {(t, v) for t in (1, 2, 4, 8, 16, 32) for v in (1, 2)}
    assert all(c.waves_per_simd == 1 and c.strips == 1 for c in tuning.CANDIDATES)
    # a batch also prices the packed chain (tile heights up to 16); the first
    # candidates are the same
    seen = {(c.tile_m, c.variant) for c in tuning.BATCH_CANDIDATES}
    assert seen == ({(t, v) for t in (1, 2, 4, 8, 16, 32) for v in (1, 2)} |
                    {(t, 3) for t in (1, 2, 4, 8, 16)})
    assert tuning.BATCH_CANDIDATES[:12] == tuning.CANDIDATES


def test_features_are_scaled_inputs():
    X = tuning.features([256, 16384, 262144], [300, 300, 300], [4, 20, 4], [0.05, 0.2, 0.1])
    assert X.shape == (3, 4) and X.dtype == np.float32
    assert np.allclose(X[:, 0], np.log2([256, 16384, 262144]) / 18)
    assert np.allclose(X[:, 2], np.log2([4, 20, 4]) / 5)
    assert np.allclose(X[:, 3], [0.2, 0.8, 0.4])
    assert (X >= 0).all() and (X <= 1.01).all()


def test_heuristic_matches_library_rule():
    # default_batch_tile_m in csrc/lcs_batch.cpp
    assert [tuning.heuristic_tile_m(m) for m in (1, 64, 65, 128, 200, 256, 300, 512, 513, 10**6)] == \
        [1, 1, 2, 2, 4, 4, 8, 8, 16, 16]


def test_regret_is_chosen_over_best():
    C = np.array([[2.0, 1.0, 4.0], [3.0, 3.0, 1.5]])
    assert np.allclose(tuning.regret(C, np.array([1, 0])), [1.0, 2.0])


def test_shape_sampler_is_deterministic_and_in_range():
    a = tuning.sample_shapes(20, 11, 256, 16384, (4, 20))
    b = tuning.sample_shapes(20, 11, 256, 16384, (4, 20))
    assert a == b
    assert all(256 <= s["m"] <= 16384 and 256 <= s["n"] <= 16384 and s["alphabet"] in (4, 20) for s in a)
    assert all(0.05 <= s["rate"] <= 0.20 for s in a)


def test_committed_timing_tables_are_well_formed():
    for mode in (1, 0):
        t = tuning.load_table(mode)
        assert t["mode"] == mode
        cands = tuning.MODE_CANDIDATES[mode]
        assert t["candidates"] == [[c.tile_m, c.waves_per_simd, c.variant, c.strips] for c in cands]
        X, T, M, C = tuning.table_arrays(t)
        assert X.shape == (len(t["rows"]), 4) and C.shape == (len(t["rows"]), len(cands))
        assert M.all() and np.isfinite(C).all() and (C > 0).all()


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
