"""CPU emulation of the whole-GPU steady ant's index arithmetic
(platform-independent-lcs_amd/csrc/braid_kernels.hip, the "big cores" kernels:
big_split_count/scatter, big_lift_kernel, big_hist_kernel, block_search, sweep_block,
big_walk_kernel), step for step in Python, checked against the oracle's
steady ant (orc_steady_ant, braid_multiplication.hpp:204-270).  The leaves
(one workgroup each on the GPU) are taken from the oracle here; the GPU tests
check the leaf kernel itself.  Small line and block constants make small
cores exercise what only large cores meet on the GPU: several histogram
blocks per node, blocks of 512 entries, boxes wider than 64."""
import numpy as np
import pytest

import oracle

LO = 1 << 31


def counted31(e, r):
    return (((e - r) & 0xFFFFFFFF) >> 31) & 1


def counted16(e, r):
    return (((e - r) & 0xFFFFFFFF) >> 15) & 1


def w_log(e):
    wl = 1 + (e + 1) // 2
    return min(max(wl, 9), 13)


def depth(L, d, box_g, max_lines, w_fixed):
    s = 1 << (L - d)
    nb = min(s // box_g, max_lines)
    glog = (s // nb).bit_length() - 1
    w = min(s, w_fixed if w_fixed else 1 << w_log(L - d))
    return s, nb, glog, w, s // w


def partition(key, s, bit, cnt, tile=2048, next_bit=None):
    """big_split_scatter_kernel: stable partition of every segment of size s
    by `bit`, zeros first; par[pos] = old position.  cnt holds this depth's
    ones per tile; a block sums its segment's tiles itself.  With next_bit,
    it also returns the next depth's tile counts as the scatter accumulates
    them: each block counts the next bit of the keys it places per
    destination tile, its zeros and its ones each running into at most two
    tiles from their first destinations."""
    K = key.size
    new = np.empty_like(key)
    par = np.empty(K, np.int64)
    nxt = np.zeros(K // tile, np.int64)
    for blk in range(K // tile):
        seg = (blk * tile) & ~(s - 1)
        t0, nt = seg // tile, s // tile
        before_tiles = int(sum(cnt[t0:t0 + (blk - t0)]))
        ones_total = int(sum(cnt[t0:t0 + nt]))
        zeros_total = s - ones_total
        z0, o0 = blk * tile - before_tiles, seg + zeros_total + before_tiles
        zt, ot = z0 // tile, o0 // tile
        runs = {("z", 0): 0, ("z", 1): 0, ("o", 0): 0, ("o", 1): 0}
        before = before_tiles
        for x in range(blk * tile, (blk + 1) * tile):
            b = (int(key[x]) >> bit) & 1
            pos = seg + zeros_total + before if b else seg + (x - seg - before)
            before += b
            new[pos] = key[x]
            par[pos] = x
            if next_bit is not None and (int(key[x]) >> next_bit) & 1:
                idx = pos // tile - (ot if b else zt)
                assert idx in (0, 1)
                runs[("o" if b else "z", idx)] += 1
        if next_bit is not None:
            for (kind, idx), v in runs.items():
                if v:
                    nxt[(ot if kind == "o" else zt) + idx] += v
    return new, par, nxt


def tile_counts(key, bit, tile=2048):
    """big_split_count_kernel (depth 0): ones of `bit` per tile."""
    return np.array([int(((key[b * tile:(b + 1) * tile] >> bit) & 1).sum()) for b in range(key.size // tile)])


def block_search(Hj, Ht, nsegs, Yj):
    """Rounds of 64 blocks: the first block where Z + its own count of z_j
    entries reaches Y_j (the lanes' inclusive scan plus a carry)."""
    Z = 0
    for g0 in range(0, nsegs, 64):
        cs = []
        for g in range(g0, min(g0 + 64, nsegs)):
            P, T = Hj[g], Ht[g]
            cs.append((P & 0xFFFF) + (T >> 16) - (P >> 16))
        incl = np.cumsum(cs)
        hit = np.nonzero(Z + incl >= Yj)[0]
        if hit.size:
            l = int(hit[0])
            return g0 + l, Z + int(incl[l] - cs[l])
        Z += int(incl[-1])
    return nsegs - 1, Z


def emulate(P, Q, k, L, L0, box_g=64, max_lines=8192, w_fixed=None):
    K = 1 << L
    d0 = L - L0
    keyP = np.array([P[x] if x < k else x for x in range(K)], np.int64)
    keyQ = np.arange(K, dtype=np.int64)
    for r in range(k):
        keyQ[Q[r]] = r
    parP, parQ = [], []
    cP, cQ = tile_counts(keyP, L - 1), tile_counts(keyQ, L - 1)
    for d in range(d0):
        nb = L - d - 2 if d + 1 < d0 else None
        keyP, pp, nP = partition(keyP, K >> d, L - d - 1, cP, next_bit=nb)
        keyQ, pq, nQ = partition(keyQ, K >> d, L - d - 1, cQ, next_bit=nb)
        if nb is not None:
            # the accumulated counts are the next depth's tile counts
            assert np.array_equal(nP, tile_counts(keyP, nb)) and np.array_equal(nQ, tile_counts(keyQ, nb))
            cP, cQ = nP, nQ
        parP.append(pp)
        parQ.append(pq)
    s0 = 1 << L0
    R = np.empty(K, np.int64)
    for v in range(K >> L0):
        kp = keyP[v * s0:(v + 1) * s0] & (s0 - 1)
        kq = keyQ[v * s0:(v + 1) * s0] & (s0 - 1)   # column -> shared row
        qr2c = np.empty(s0, np.int32)
        qr2c[kq] = np.arange(s0)
        R[v * s0:(v + 1) * s0] = oracle.steady_ant(kp.astype(np.int32), qr2c)
    for d in range(d0 - 1, -1, -1):
        s, nb, glog, w, nsegs = depth(L, d, box_g, max_lines, w_fixed)
        g = 1 << glog
        h = s >> 1
        nodes = K // s
        # big_lift_kernel
        U = np.zeros(K, np.int64)
        CI = np.zeros(K, np.int64)
        Rn = np.zeros(K, np.int64)
        for x in range(K):
            p_pos = int(parP[d][x])
            q_pos = int(parQ[d][(x & ~(h - 1)) + int(R[x])])
            v0 = p_pos & ~(s - 1)
            lo = 0 if (x & h) else LO
            U[p_pos] = (q_pos - v0) | lo
            CI[v0 + q_pos - v0] = (p_pos - v0) | lo
            Rn[p_pos] = q_pos - v0
        # big_hist_kernel: per (y, node) an nb x nsegs prefix matrix, then totals
        H = {}
        for y, E in ((0, CI), (1, U)):
            for node in range(nodes):
                M = np.zeros((nb + 1, nsegs), np.int64)
                for gg in range(nsegs):
                    hh = np.zeros(nb, np.int64)
                    for i in range(w):
                        e = int(E[node * s + gg * w + i])
                        bk = (s - 1 - (e & 0x7FFFFFFF)) >> glog
                        hh[bk] += 1 if (e & LO) else 0x10000
                    M[:nb, gg] = np.concatenate([[0], np.cumsum(hh)[:-1]])
                    M[nb, gg] = hh.sum()
                H[(y, node)] = M
        # big_sweep_kernel: block_search + sweep_block
        T = {}
        for y, E in ((0, CI), (1, U)):
            for node in range(nodes):
                M = H[(y, node)]
                for j in range(nb):
                    Yj = int(sum(int(v) & 0xFFFF for v in M[j])) if j else 0
                    if Yj == 0:
                        T[(y, node, j)] = 0
                        continue
                    seg, Z = block_search(M[j], M[nb], nsegs, Yj)
                    need = Yj - Z
                    base = node * s + seg * w
                    for r in range(w):
                        e = int(E[base + r])
                        bk = (s - 1 - (e & 0x7FFFFFFF)) >> glog
                        if ((bk < j) if (e & LO) else (bk >= j)):
                            need -= 1
                            if need == 0:
                                T[(y, node, j)] = seg * w + r + 1
                                break
                    assert need == 0
        # big_walk_kernel: one lane per piece, 16-bit entries relative to the box
        for node in range(nodes):
            v0 = node * s
            for q in range(2 * nb - 1):
                if q < nb:
                    Dr = s - q * g
                    dc = 0 if q == 0 else min(max(T[(0, node, q)], 0), s)
                    dr, r_lo = Dr - 1, Dr - g
                    cb = 0 if dc == 0 else (dc - 1) >> glog
                else:
                    i = q - nb + 1
                    Tc = min(max(T[(1, node, nb - i)], 0), s)
                    if Tc == 0:
                        continue
                    dr, dc = Tc - 1, i * g + 1
                    r_lo = s - (((s - 1 - dr) >> glog) + 1) * g
                    cb = i
                c_lo = cb * g + 1 if cb else 0
                c_hi = s + 1 if cb == nb - 1 else (cb + 1) * g + 1
                nc = c_hi - c_lo + 1
                assert nc <= g + 2

                def box_entry(e, base, top):
                    return (0x8000 if (e & LO) else 0) | (min(max(e & 0x7FFFFFFF, base), top) - base)

                bu = [box_entry(int(U[v0 + min(max(r_lo - 1 + i, 0), s - 1)]), c_lo - 1, c_hi) for i in range(g + 1)]
                bc = [box_entry(int(CI[v0 + min(max(c_lo - 1 + i, 0), s - 1)]), r_lo - 1, r_lo + g) for i in range(nc)]
                box = bu + bc
                Rr, X, Xs = dr - r_lo + 1, dc - c_lo + 1, s - c_lo + 1
                assert 1 <= Rr <= g and 1 <= X < nc
                # the same entry values as the node-level (absolute) form
                if q < nb:
                    assert counted16(bu[Rr], X) == counted31(int(U[v0 + dr]), dc)
                    delta = -counted16(bu[Rr], X)
                else:
                    assert counted16(bc[0], Rr) == counted31(int(CI[v0 + i * g]), dr)
                    delta = -1 + counted16(bc[0], Rr)
                cl, cc, ud, ur = bc[X - 1], bc[X], bu[Rr], bu[Rr - 1]
                while Rr >= 1 and X < nc:
                    d_left = delta - counted16(cl, Rr)
                    d_down = delta + counted16(ud, X)
                    if X + c_lo > 1 and d_left < 0 and d_down > 0:
                        Rn[v0 + Rr + r_lo - 1] = X + c_lo - 2
                    right = delta < 0 and X < Xs
                    d_right = counted16(cc, Rr)
                    d_up = counted16(ur, X) if Rr + r_lo > 1 else 0
                    delta = delta + d_right if right else delta - d_up
                    X += 1 if right else 0
                    Rr -= 0 if right else 1
                    cl = cc if right else cl
                    ud = ud if right else ur
                    nx = box[g + 1 + min(X, nc - 1)] if right else box[max(Rr - 1, 0)]
                    cc = nx if right else cc
                    ur = ur if right else nx
        R = Rn
    return R[:k].astype(np.int32)


CASES = [
    # (k, L, L0, box side, max lines, fixed block width)
    (4096, 12, 10, 64, 8192, None),     # the library's constants: one block per node
    (4000, 12, 11, 64, 8192, None),     # padded core, leaves of 2^11
    (6000, 13, 11, 32, 8192, None),     # boxes of 32
    (4096, 12, 10, 64, 16, 512),        # wide boxes (g = 256 at the top) and 8 blocks per node
    (8192, 13, 10, 16, 8192, 512),      # boxes of 16, 16 blocks per node
    (16384, 14, 10, 32, 8192, None),    # four device-wide depths, 8 partition tiles per segment
]


@pytest.mark.parametrize("k,L,L0,box_g,max_lines,w", CASES)
def test_big_core_random(k, L, L0, box_g, max_lines, w):
    rng = np.random.default_rng(k + L0 + box_g)
    P = rng.permutation(k).astype(np.int32)
    Q = rng.permutation(k).astype(np.int32)
    assert np.array_equal(emulate(P, Q, k, L, L0, box_g, max_lines, w), oracle.steady_ant(P, Q))


def test_big_core_structured():
    """Identity, reversal, a shift and few swaps: paths hugging the box edges."""
    k = 4096
    rng = np.random.default_rng(1)
    ident = np.arange(k, dtype=np.int32)
    rev = ident[::-1].copy()
    shift = np.roll(ident, 123).astype(np.int32)
    near = ident.copy()
    for _ in range(50):
        i, j = rng.integers(0, k, 2)
        near[i], near[j] = near[j], near[i]
    for P, Q in ((ident, rev), (rev, rev), (shift, near), (rev, ident)):
        for box_g, max_lines, w in ((64, 8192, None), (64, 16, 512)):
            assert np.array_equal(emulate(P, Q, k, 12, 10, box_g, max_lines, w), oracle.steady_ant(P, Q))


def test_block_width_bounds():
    """Blocks are 512 .. 8192 entries, at most 1024 per node up to 2^22 points
    (what big_sweep_kernel's block search holds in registers)."""
    for e in range(11, 23):
        w = min(1 << e, 1 << w_log(e))
        assert 512 <= w <= 8192 and (1 << e) // w <= 1024
