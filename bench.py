#!/usr/bin/env python3
"""Benchmark of the semi-local LCS seaweed hot path on one MI355X.

Default workload (one "step" = one full comb of the pair, inputs resident in
HBM): m = n = 2^20, A uniform over alphabet 4, B = A with 10% substitutions
and 2% single-symbol indels (seeded), output the int32 permutation of 2^21
strands.  Prints ONE JSON line with GCUPS, the integer-VALU roofline of the
comb kernel and a CPU baseline measured in the same run on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--tile-m T]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "platform-independent-lcs_amd")]

import numpy as np  # noqa: E402

import lcs_amd  # noqa: E402
from lcs_amd import hip as H  # noqa: E402

# MI355X: 256 CUs x 4 SIMD x 16 lanes = 64 int32 lane-ops per CU per clock,
# 2.4 GHz peak engine clock.  A cell is 4 VALU ops (cmp, min, cndmask, xor3).
PEAK_CLOCK_HZ = 2.4e9
OPS_PER_CELL = 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=20, help="m = n = 2^log2n")
    ap.add_argument("--tile-m", type=int, default=0)
    ap.add_argument("--waves-per-simd", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0, help="0 auto, 1 register chain, 2 LDS band groups")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU baseline sample length")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check", action="store_true", help="cross-check two schedules bit-for-bit")
    return ap.parse_args()


def cpu_baseline(a, b, target_s: float):
    """The oracle's multithreaded comb (a port of the reference's block
    wavefront) on a full-height strip of the same pair, scaled to GCUPS."""
    import oracle  # test infrastructure: the CPU baseline only
    threads = min(16, len(os.sched_getaffinity(0)))
    m = len(a)
    # calibrate on a small strip, then size the sample for ~target_s seconds
    cols = 256
    t0 = time.perf_counter()
    oracle.comb(a, b[:cols], order="mt", threads=threads)
    rate = m * cols / max(time.perf_counter() - t0, 1e-6)
    cols = int(min(len(b), max(512, rate * target_s / m)))
    t0 = time.perf_counter()
    oracle.comb(a, b[:cols], order="mt", threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round(m * cols / dt / 1e9, 4), "unit": "GCUPS", "cores": threads, "kind": "port",
            "sample": f"m={m} x first {cols} columns of the same pair ({m * cols:.3g} cells, {dt:.1f} s), "
                      f"oracle orc_comb_mt block wavefront, {threads} threads"}


def read_traffic(path: str):
    """HBM bytes per launch of the comb kernel from a committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE, KiB per dispatch, scaled by the
    copy-kernel calibration factor stated in that file)."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU sharding is a design only)", file=sys.stderr)
        sys.exit(2)
    size = 1 << args.log2n
    wl = lcs_amd.workload_single(f"mutated2^{args.log2n}", seed=1)
    a, b = wl.a, wl.b
    m, n = a.size, b.size
    ctx = lcs_amd.Context(0)
    cus, _ = ctx.info()
    sched = lcs_amd.schedule(tile_m=args.tile_m, waves_per_simd=args.waves_per_simd, variant=args.variant)
    da, db = H.DeviceArray.from_numpy(a), H.DeviceArray.from_numpy(b)
    dh, dv = H.DeviceArray(m), H.DeviceArray(n)
    L = lcs_amd.lib()

    def step():
        rc = L.lcs_comb_device(ctx.handle, da.ptr, m, db.ptr, n, sched, dh.ptr, dv.ptr, None)
        if rc:
            raise lcs_amd.LcsError("lcs_comb_device", rc)
        H.synchronize()
        return ctx.info()[1]

    for _ in range(args.warmup):
        step()
    H.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        kernel_ms.append(step())
    H.synchronize()
    wall = time.perf_counter() - t0
    h, v = dh.numpy(), dv.numpy()
    # size-independent checks: the output is a bijection on [0, m+n)
    r2c, c2r = lcs_amd.from_strands(h, v)
    assert (r2c >= 0).all() and (c2r >= 0).all(), "output strands are not a permutation"
    llcs = int((v < m).sum())
    checked = None
    if args.check:
        alt = lcs_amd.schedule(tile_m=8 if (args.tile_m or 16) != 8 else 4)
        h2, v2, _ = ctx.comb(a, b, alt)
        checked = bool(np.array_equal(h, h2) and np.array_equal(v, v2))
    cells = float(m) * float(n)
    ms_step = wall / args.steps * 1e3
    gcups = cells * args.steps / wall / 1e9
    k_ms = float(np.mean(kernel_ms))
    peak_ops = cus * 64 * PEAK_CLOCK_HZ
    achieved_ops = OPS_PER_CELL * cells / (k_ms / 1e3)
    traffic = read_traffic(os.path.join(ROOT, "profiles", "comb_chain_mutated2^20_pmc.txt"))
    out = {
        "metric": "GCUPS at m=n=2^20 (seaweed permutation, bit-exact)",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded mt19937_64: uniform alphabet-4 A, B = A + 10% subs + 2% indels)",
        "config": {"workload": f"single pair m=n=2^{args.log2n}, tiled band-chain comb", "m": m, "n": n,
                   "alphabet": 4, "tile_m": args.tile_m or None, "variant": args.variant or None, "llcs": llcs},
        "roofline": {"bound": "valu", "achieved": round(achieved_ops / 1e12, 3), "peak": round(peak_ops / 1e12, 3),
                     "unit": "TOP/s (int32 lane-ops, 4 per cell)", "frac": round(achieved_ops / peak_ops, 4),
                     "traffic": traffic, "kernel_ms": round(k_ms, 3),
                     "kernel_gcups": round(cells / (k_ms / 1e3) / 1e9, 2)},
    }
    if checked is not None:
        out["cross_schedule_bit_exact"] = checked
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a, b, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
