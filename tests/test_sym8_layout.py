"""Index arithmetic of the register chain's byte symbols, checked on the CPU.

comb_sym8_kernel writes four copies of b (copy j holds at byte x the symbol
of column x + j - 128, zero outside [0, n)); a band lane L of a strip whose
first column is col0 reads, for the chunk starting at step t, C/4 = 8 words
of copy j = x0 % 4 from word (x0 - j) / 4 + t / 4, x0 = 128 + col0 - L, and
takes byte s % 4 of word s / 4 as the symbol of step t + s.  This restates
those formulas (comb_kernels.hip: comb_sym8_kernel and band_pass) and checks
for every lane, every chunk up to the last prefetch, several strip offsets
and widths that each read stays inside its copy and that each byte a
combing lane uses is the symbol of the column it combs.  It runs no GPU
code; the GPU parity tests compare the kernels themselves with the oracle.
"""
import numpy as np
import pytest

PAD, C, WAVE = 128, 32, 64


def stride_words(n):
    return ((n + 2 * PAD + 15) // 16) * 4


def copies(b):
    n = len(b)
    st = stride_words(n)
    out = np.zeros((4, st * 4), np.uint8)
    for j in range(4):
        x = np.arange(st * 4)
        c = x + j - PAD
        ok = (c >= 0) & (c < n)
        out[j, ok] = (b[c[ok]] & 0xFF).astype(np.uint8)
    return out, st


@pytest.mark.parametrize("n,strips", [(1, 1), (63, 1), (64, 1), (97, 1), (1000, 1), (5003, 2), (5003, 3),
                                      (5003, 7), (4096, 4), (777, 5)])
def test_lane_windows_stay_inside_and_pick_their_columns(n, strips):
    rng = np.random.default_rng(n * 10 + strips)
    b = rng.integers(0, 256, n).astype(np.int64)
    cp, st = copies(b)
    assert st * 4 >= n + 2 * PAD and st % 4 == 0
    W = (n + strips - 1) // strips
    for s0 in range(0, n, W):
        ns = min(W, n - s0)
        nchunks = (ns + WAVE + C - 1) // C
        lanes = np.arange(WAVE)
        x0 = PAD + s0 - lanes
        j = x0 & 3
        word0 = (x0 - j) // 4
        assert (word0 >= 0).all()
        for k in range(nchunks + 1):        # the last chunk prefetches chunk nchunks
            t = k * C
            first = word0 + t // 4
            last = first + C // 4 - 1
            assert (first >= 0).all() and (last < st).all(), (n, strips, s0, k)
            if k == nchunks:
                continue
            for L in range(WAVE):
                words = cp[j[L], 4 * first[L]: 4 * (last[L] + 1)]
                for s in range(C):
                    col = t + s - L
                    if 0 <= col < ns:
                        assert words[s] == b[s0 + col], (n, strips, s0, k, L, s)


def test_copies_zero_outside_the_string():
    b = np.arange(1, 11, dtype=np.int64)
    cp, st = copies(b)
    for j in range(4):
        nz = np.nonzero(cp[j])[0]
        assert nz.min() == PAD - j and nz.max() == PAD - j + 9
