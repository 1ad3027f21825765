"""Parity at the full sizes of the benchmark workloads (bit-exact).

The oracle's vectorised stripe comb (pinned against the sequential comb in
test_oracle.py) recombs each workload on the host cores, so the GPU's whole
output is compared, not a property of it:
  - the single pair m = n = 2^20 (B = A + 10% substitutions + 2% indels):
    the full permutation of 2^21 strands, plus sampled semi-local windows of
    it against the word-parallel LCS of the corresponding substrings;
  - the Zipf pair m = 2^14, n = 2^22 (alphabet 26, 64 planted copies):
    strips merged by the GPU steady ant, against the whole comb;
  - the 10,000-pair batch (lengths log-uniform in [256, 16384]), every pair;
  - served requests with lengths near 262,144 (alphabets 4 and 20) in a mix
    with short ones."""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import lcs_amd
import oracle
from conftest import cpu_threads

pytestmark = pytest.mark.gpu


def _oracle(a, b, threads=None):
    return oracle.comb(a, b, order="simd", threads=threads or cpu_threads())


def _log(msg):
    print(msg, flush=True)


def test_mutated_2p20_full_permutation(ctx):
    w = lcs_amd.workload_single("mutated2^20", seed=1)
    m, n = w.a.size, w.b.size
    _log(f"mutated2^20: GPU comb m={m} n={n}")
    r2c, c2r, llcs = ctx.comb_permutation(w.a, w.b)
    t0 = time.perf_counter()
    ho, vo = _oracle(w.a, w.b)
    _log(f"mutated2^20: oracle simd {time.perf_counter() - t0:.1f} s, {cpu_threads()} threads")
    ro, co = oracle.from_strands(ho, vo)
    assert np.array_equal(c2r, co), "col_to_row differs from the oracle"
    assert np.array_equal(r2c, ro), "row_to_col differs from the oracle"
    assert llcs == oracle.llcs_from_strands(vo, m)
    # semi-local windows read from the GPU permutation against the
    # word-parallel LCS of the substrings themselves (an independent algorithm)
    rng = np.random.default_rng(20)
    q0, q1 = [], []
    for _ in range(48):   # string-substring LCS(a, b[i:j)), windows up to 4096 columns
        i = int(rng.integers(0, n - 4096))
        q0.append(i)
        q1.append(i + int(rng.integers(0, 4097)))
    got = ctx.semilocal(r2c, m, n, lcs_amd.SEMI_STRING_SUBSTRING, q0, q1)
    for i, j, g in zip(q0, q1, got):
        assert g == oracle.llcs_bitpar(w.a, w.b[i:j]), (i, j)
    p0, p1 = [], []
    for _ in range(16):   # prefix-suffix LCS(a[:p), b[q:)), suffixes up to 4096
        p0.append(int(rng.integers(0, m + 1)))
        p1.append(n - int(rng.integers(0, 4097)))
    got = ctx.semilocal(r2c, m, n, lcs_amd.SEMI_PREFIX_SUFFIX, p0, p1)
    for p, q, g in zip(p0, p1, got):
        assert g == oracle.llcs_bitpar(w.a[:p], w.b[q:]), (p, q)


def test_zipf_planted_full_pair(ctx):
    w = lcs_amd.workload_single("zipf_planted", seed=1)
    _log(f"zipf_planted: m={w.a.size} n={w.b.size}, default schedule (strips + steady-ant merges)")
    h, v, llcs = ctx.comb(w.a, w.b)
    ho, vo = _oracle(w.a, w.b)
    assert np.array_equal(h, ho) and np.array_equal(v, vo)
    assert llcs == oracle.llcs_from_strands(vo, w.a.size)


def test_batch_10000_pairs_every_pair(ctx):
    pairs = lcs_amd.workload_batch(10000, seed=7)
    a_list, b_list = [p.a for p in pairs], [p.b for p in pairs]
    _log("batch10000: GPU batch (heuristic schedules, hipGraph)")
    hs, vs = ctx.batch_comb(a_list, b_list, use_graph=True)
    with ThreadPoolExecutor(cpu_threads()) as ex:   # one thread per pair, pairs side by side
        ref = list(ex.map(lambda p: _oracle(p.a, p.b, threads=1), pairs))
    bad = [i for i, (h, v, (ho, vo)) in enumerate(zip(hs, vs, ref))
           if not (np.array_equal(h, ho) and np.array_equal(v, vo))]
    assert not bad, f"{len(bad)} pairs differ, first {bad[:5]}"
    # per-pair schedules across every tile height and both variants: same bits
    scheds = [lcs_amd.schedule(tile_m=[1, 2, 4, 8, 16, 32][i % 6], variant=1 + (i // 6) % 2)
              for i in range(len(pairs))]
    _log("batch10000: GPU batch (mixed explicit schedules)")
    hs2, vs2 = ctx.batch_comb(a_list, b_list, scheds=scheds, use_graph=True)
    bad = [i for i, (h, v, (ho, vo)) in enumerate(zip(hs2, vs2, ref))
           if not (np.array_equal(h, ho) and np.array_equal(v, vo))]
    assert not bad, f"{len(bad)} pairs differ with explicit schedules, first {bad[:5]}"


def test_served_long_requests_in_a_mix(ctx):
    long = lcs_amd.workload_batch(6, seed=31, lo=200000, hi=262144, alphabet=(4, 20))
    short = lcs_amd.workload_batch(10, seed=32, lo=1024, hi=8192, alphabet=(4, 20))
    mix = [x for pair in zip(long, short[:6]) for x in pair] + short[6:]
    srv = lcs_amd.Server(ctx, hbm_budget=1 << 30, max_inflight=4)
    try:
        _log(f"served: {len(mix)} requests, {len(long)} of length >= 200000")
        ids = [srv.submit(p.a, p.b) for p in mix]
        outs = [srv.wait(i, timeout_ms=120000) for i in ids]
        assert srv.stats()["status"] == 0
    finally:
        srv.close()
    for p, (h, v) in zip(mix, outs):
        ho, vo = _oracle(p.a, p.b)
        assert np.array_equal(h, ho) and np.array_equal(v, vo), (p.a.size, p.b.size, p.alphabet)
