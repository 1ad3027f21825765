"""Schedule autotuning and the learned cost model.

The reference picks its kernel shape by name at compile time
(src/algorithm_registry.cpp:127-142) and swept shapes offline in notebooks
(tests/test_prepare.ipynb).  Here the library's autotuner
(lcs_autotune_measure) times every candidate schedule on synthetic pairs of
a shape; a table of such timings trains the cost model (lcs_mlp_*: an fp32
MLP 4 -> 64 -> 64 -> one cost per candidate, trained and evaluated by the library's
own HIP kernels) to predict log2 cost per candidate from

    (log2 m, log2 n, log2 alphabet, mutation rate)        (scaled, see features)

and the cheapest prediction is the schedule.  Exhaustive autotuning of the
shape is the fallback and the yardstick of a choice (its regret: chosen cost /
best cost).  Schedules change speed only; every candidate computes the same
strands.

    python -m lcs_amd.tuning record --mode 1 --shapes 48 --out table.json   (on a GPU)
    python -m lcs_amd.tuning report --table table.json                     (on a GPU)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np

from . import Context, Mlp, Schedule, gen_unit, schedule

VARIANT_CHAIN, VARIANT_CHAIN_LDS = 1, 2
TILES = (1, 2, 4, 8, 16, 32)
# every (tile height, variant) pair, one wave per SIMD, no column strips
CANDIDATES = [schedule(tile_m=t, waves_per_simd=1, variant=v, strips=1) for v in (VARIANT_CHAIN, VARIANT_CHAIN_LDS)
              for t in TILES]
# a batch also prices the 16-bit packed chain (two pairs per wave); for a
# single pair it is the register chain, so single-pair tables leave it out.
# Its 32-row tile is left out: on every recorded shape it cost 1.0x to 10x
# the 8-row tile, yet a model fitted to those prices still sent some pairs
# of the 10,000-pair batch there, and that tile class alone held the batch
# at 75 ms against 15 ms without it.
VARIANT_PACKED16 = 3
PACKED_TILES = (1, 2, 4, 8, 16)
BATCH_CANDIDATES = CANDIDATES + [schedule(tile_m=t, waves_per_simd=1, variant=VARIANT_PACKED16, strips=1)
                                 for t in PACKED_TILES]
MODE_CANDIDATES = {0: CANDIDATES, 1: BATCH_CANDIDATES}


def table_candidates(table: dict):
    """The candidate schedules a table's cost columns belong to."""
    return [Schedule(*c) for c in table["candidates"]]
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
TABLES = {1: os.path.join(DATA_DIR, "autotune_batch_mi355x.json"),
          0: os.path.join(DATA_DIR, "autotune_single_mi355x.json")}


def features(m, n, alphabet, rate) -> np.ndarray:
    """Cost-model inputs, each scaled to about [0, 1]."""
    m, n, alphabet, rate = (np.atleast_1d(np.asarray(x, np.float64)) for x in (m, n, alphabet, rate))
    return np.stack([np.log2(np.maximum(m, 1)) / 18.0, np.log2(np.maximum(n, 1)) / 18.0,
                     np.log2(np.maximum(alphabet, 2)) / 5.0, rate * 4.0], axis=1).astype(np.float32)


def heuristic_tile_m(m: int) -> int:
    """The library's built-in batch choice (default_batch_tile_m): one band of
    64 rows per short string, up to 16 rows per lane for long ones."""
    for limit, tm in ((64, 1), (128, 2), (256, 4), (512, 8)):
        if m <= limit:
            return tm
    return 16


def heuristic_index(m: int) -> int:
    return TILES.index(heuristic_tile_m(m))   # register chain, as the library does


def sample_shapes(count: int, seed: int, lo: int, hi: int, alphabets=(4,), rate_lo=0.05, rate_hi=0.20):
    """Shapes drawn like the workloads: m, n log-uniform in [lo, hi]."""
    out = []
    for i in range(count):
        u = [gen_unit(seed, 4 * i + q) for q in range(4)]
        m = int(min(hi, max(lo, math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u[0]))))
        n = int(min(hi, max(lo, math.exp(math.log(lo) + (math.log(hi) - math.log(lo)) * u[1]))))
        alpha = alphabets[int(u[2] * len(alphabets)) % len(alphabets)]
        out.append({"m": m, "n": n, "alphabet": int(alpha), "rate": round(rate_lo + (rate_hi - rate_lo) * u[3], 4)})
    return out


def record(ctx: Context, shapes, mode: int = 1, reps: int = 1, seed: int = 1, log=None) -> dict:
    """Runs the autotuner on every shape; returns the timing table."""
    rows = []
    t_start = time.perf_counter()
    for i, sh in enumerate(shapes):
        t0 = time.perf_counter()
        cost = ctx.autotune(sh["m"], sh["n"], sh["alphabet"], sh["rate"], MODE_CANDIDATES[mode], mode=mode, reps=reps,
                            seed=seed + i)
        rows.append(dict(sh, cost_ms=[float(c) for c in cost], tune_s=round(time.perf_counter() - t0, 3)))
        if log:
            log(f"shape {i + 1}/{len(shapes)} m={sh['m']} n={sh['n']} a={sh['alphabet']} "
                f"best={int(np.argmin(cost))} {time.perf_counter() - t0:.2f}s")
    return {"mode": mode,
            "candidates": [[c.tile_m, c.waves_per_simd, c.variant, c.strips] for c in MODE_CANDIDATES[mode]],
            "rows": rows, "device": ctx.device_name(), "seconds": round(time.perf_counter() - t_start, 1)}


def table_arrays(table: dict):
    rows = table["rows"]
    X = features([r["m"] for r in rows], [r["n"] for r in rows], [r["alphabet"] for r in rows],
                 [r["rate"] for r in rows])
    C = np.array([r["cost_ms"] for r in rows], np.float64)
    ok = np.isfinite(C) & (C > 0)
    T = np.where(ok, np.log2(np.where(ok, C, 1.0)), 0.0).astype(np.float32)
    return X, T, ok.astype(np.float32), C


def train(ctx: Context, table: dict, rows=None, steps: int = 4000, lr: float = 3e-3, seed: int = 1,
          hidden: int = 64) -> Mlp:
    """Trains the cost model on the GPU on (a subset of) the table's rows."""
    X, T, M, _ = table_arrays(table)
    if rows is not None:
        X, T, M = X[rows], T[rows], M[rows]
    mlp = Mlp(ctx, n_in=X.shape[1], hidden=hidden, n_out=T.shape[1], seed=seed)
    mlp.train(X, T, M, steps=steps, lr=lr)
    return mlp


def load_table(mode: int = 1) -> dict:
    with open(TABLES[mode]) as f:
        return json.load(f)


def best_rows(table: dict):
    """(m, n, best Schedule) of every measured shape of a table."""
    cands = [Schedule(*c) for c in table["candidates"]]
    out = []
    for r in table["rows"]:
        c = np.array(r["cost_ms"], np.float64)
        c = np.where(np.isfinite(c) & (c > 0), c, np.inf)
        if np.isfinite(c).any():
            out.append((int(r["m"]), int(r["n"]), cands[int(np.argmin(c))]))
    return out


def load_into(ctx: Context, modes=(0, 1), model: bool = False, steps: int = 4000):
    """Loads the committed autotune tables of this device into the context's
    schedule selection (step 1 of lcs_ctx_choose_schedule); with model=True
    also trains a cost model per mode on the GPU and attaches it (step 2).
    Returns the attached models (close them after the context stops using
    them)."""
    models = {}
    for mode in modes:
        if not os.path.exists(TABLES[mode]):
            continue
        table = load_table(mode)
        for m, n, s in best_rows(table):
            ctx.add_tuned(mode, m, n, s)
        if model:
            mlp = train(ctx, table, steps=steps)
            ctx.set_cost_model(mode, mlp, [Schedule(*c) for c in table["candidates"]])
            models[mode] = mlp
    return models


def regret(C: np.ndarray, choice: np.ndarray) -> np.ndarray:
    """Cost of the chosen schedule over the best measured one, per row."""
    best = np.min(np.where(np.isfinite(C), C, np.inf), axis=1)
    return C[np.arange(len(C)), choice] / best


def evaluate(ctx: Context, table: dict, folds: int = 4, steps: int = 4000, seed: int = 1) -> dict:
    """k-fold held-out regret of the model against the built-in heuristic."""
    X, _, _, C = table_arrays(table)
    rows = np.arange(len(C))
    rng = np.random.default_rng(seed)
    rng.shuffle(rows)
    model_choice = np.zeros(len(C), np.int64)
    for f in range(folds):
        test = rows[f::folds]
        train_rows = np.setdiff1d(rows, test)
        mlp = train(ctx, table, rows=train_rows, steps=steps, seed=seed + f)
        model_choice[test] = mlp.choose(X[test])
        mlp.close()
    heur = np.array([heuristic_index(r["m"]) for r in table["rows"]])
    r_model, r_heur = regret(C, model_choice), regret(C, heur)
    tune_s = float(np.mean([r.get("tune_s", 0.0) for r in table["rows"]]))
    return {"rows": len(C), "model_regret_mean": float(np.mean(r_model)), "model_regret_max": float(np.max(r_model)),
            "heuristic_regret_mean": float(np.mean(r_heur)), "heuristic_regret_max": float(np.max(r_heur)),
            "exhaustive_tune_s_per_shape": tune_s}


def choose_schedules(mlp: Mlp, m, n, alphabet, rate, candidates=None):
    """Model's schedule per pair (a list of Schedule); `candidates` are the
    schedules of the model's outputs (its table's, CANDIDATES by default)."""
    cands = candidates if candidates is not None else CANDIDATES
    idx = mlp.choose(features(m, n, alphabet, rate))
    return [cands[int(i)] for i in idx], idx


def _main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("record", help="time every candidate schedule on sampled shapes")
    r.add_argument("--mode", type=int, default=1, help="1: batched cost per pair, 0: single-pair kernel time")
    r.add_argument("--shapes", type=int, default=48)
    r.add_argument("--seed", type=int, default=11)
    r.add_argument("--lo", type=int, default=256)
    r.add_argument("--hi", type=int, default=16384)
    r.add_argument("--alphabets", default="4")
    r.add_argument("--out", required=True)
    e = sub.add_parser("report", help="held-out regret of the model vs the built-in heuristic")
    e.add_argument("--table", required=True)
    args = ap.parse_args()
    ctx = Context(0)
    if args.cmd == "record":
        shapes = sample_shapes(args.shapes, args.seed, args.lo, args.hi,
                               tuple(int(x) for x in args.alphabets.split(",")))
        table = record(ctx, shapes, mode=args.mode, log=lambda s: print(s, flush=True))
        table["recorded_with"] = (f"python -m lcs_amd.tuning record --mode {args.mode} --shapes {args.shapes} "
                                  f"--seed {args.seed} --lo {args.lo} --hi {args.hi} --alphabets {args.alphabets}")
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        with open(args.out, "w") as f:
            json.dump(table, f, indent=1)
        print(f"wrote {args.out} ({len(shapes)} shapes, {table['seconds']} s)")
    else:
        with open(args.table) as f:
            table = json.load(f)
        print(json.dumps(evaluate(ctx, table)))
    ctx.close()


if __name__ == "__main__":
    _main()
