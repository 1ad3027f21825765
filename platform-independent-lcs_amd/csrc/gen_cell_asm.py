#!/usr/bin/env python3
"""Writes cell_asm.h: hand-scheduled inline-asm blocks that comb R rows of
one column (R = 4, 8, 16) for one lane.

Per row k the block issues v_min_u32, v_cndmask_b32 (mask s[k%2]) and
v_bitop3_b32 (xor3); the compare for row k+2 is issued right after row k's
v_cndmask, so every mask is written at least two VALU issues before it is
read (the gfx950 lane-mask hazard) and the block is exactly 4R VALU issues.

cell_block{R} compares the rows' symbols with a 32-bit column symbol;
cell_block{R}_byte{q} compares them with byte q of a word of byte symbols
(SDWA operand select, zero-extended), so a lane can hold the symbols of
several columns in one register and pick one per step at no cost.
Run: python gen_cell_asm.py > cell_asm.h
"""


def block(R: int, sel=None) -> str:
    if sel is None:
        cmp = "v_cmp_eq_u32_e64 %[{s}], %[{a}], %[b]"
        name = f"cell_block{R}"
    else:
        cmp = "v_cmp_eq_u32_sdwa %[{s}], %[{a}], %[b] src0_sel:DWORD src1_sel:BYTE_" + str(sel)
        name = f"cell_block{R}_byte{sel}"
    ins = ['"' + cmp.format(s="s0", a="a0") + '\\n\\t"', '"' + cmp.format(s="s1", a="a1") + '\\n\\t"']
    for k in range(R):
        s = f"s{k % 2}"
        ins.append(f'"v_min_u32_e32 %[n{k}], %[h{k}], %[v]\\n\\t"')
        ins.append(f'"v_cndmask_b32_e64 %[n{k}], %[n{k}], %[v], %[{s}]\\n\\t"')
        ins.append(f'"v_bitop3_b32 %[v], %[h{k}], %[v], %[n{k}] bitop3:0x96\\n\\t"')
        if k + 2 < R:
            ins.append('"' + cmp.format(s=s, a=f"a{k + 2}") + '\\n\\t"')
    outs = ['[v] "+v"(v)'] + [f'[n{k}] "=&v"(hn[{k}])' for k in range(R)] + ['[s0] "=&s"(s0)', '[s1] "=&s"(s1)']
    inp = ['[b] "v"(sb)'] + [f'[a{k}] "v"(a[{k}])' for k in range(R)] + [f'[h{k}] "v"(h[{k}])' for k in range(R)]
    body = "\n        ".join(ins)
    return (f"__device__ __forceinline__ void {name}(const int32_t *a, const uint32_t *h, uint32_t *hn,"
            f" int32_t sb, uint32_t &v)\n{{\n    uint64_t s0, s1;\n    asm volatile(\n        {body}\n"
            f"        : {', '.join(outs)}\n        : {', '.join(inp)});\n}}\n")


def step_block(R: int, sel: int) -> str:
    """One whole step of the byte-symbol band loop for R rows: the two first
    compares, the strand hand-off (a copy of the conveyor Q, then v moved one
    lane up with lane 0 taking Q's lane 0, then Q moved one lane down with
    lane 63 keeping v), then the R cells on the moved strand.  The compares
    come first so the DPP read of v sits three VALU issues after the previous
    step's last v_bitop3 wrote it (gfx950 wants two), with no s_nop."""
    cmp = "v_cmp_eq_u32_sdwa %[{s}], %[{a}], %[b] src0_sel:DWORD src1_sel:BYTE_" + str(sel)
    ins = ['"' + cmp.format(s="s0", a="a0") + '\\n\\t"', '"' + cmp.format(s="s1", a="a1") + '\\n\\t"',
           '"v_mov_b32_e32 %[t], %[q]\\n\\t"',
           '"v_mov_b32_dpp %[t], %[w] wave_shr:1 row_mask:0xf bank_mask:0xf\\n\\t"',
           '"v_mov_b32_dpp %[w], %[q] wave_shl:1 row_mask:0xf bank_mask:0xf\\n\\t"']
    for k in range(R):
        s = f"s{k % 2}"
        ins.append(f'"v_min_u32_e32 %[n{k}], %[h{k}], %[t]\\n\\t"')
        ins.append(f'"v_cndmask_b32_e64 %[n{k}], %[n{k}], %[t], %[{s}]\\n\\t"')
        ins.append(f'"v_bitop3_b32 %[t], %[h{k}], %[t], %[n{k}] bitop3:0x96\\n\\t"')
        if k + 2 < R:
            ins.append('"' + cmp.format(s=s, a=f"a{k + 2}") + '\\n\\t"')
    outs = ['[w] "+v"(w)', '[t] "=&v"(t)'] + [f'[n{k}] "=&v"(hn[{k}])' for k in range(R)] + \
           ['[s0] "=&s"(s0)', '[s1] "=&s"(s1)']
    inp = ['[b] "v"(sb)', '[q] "v"(q)'] + [f'[a{k}] "v"(a[{k}])' for k in range(R)] + \
          [f'[h{k}] "v"(h[{k}])' for k in range(R)]
    body = "\n        ".join(ins)
    return (f"__device__ __forceinline__ void cell_step{R}_byte{sel}(const int32_t *a, const uint32_t *h, uint32_t *hn,"
            f" int32_t sb, uint32_t q, uint32_t &w, uint32_t &t)\n{{\n    uint64_t s0, s1;\n    asm volatile(\n"
            f"        {body}\n        : {', '.join(outs)}\n        : {', '.join(inp)});\n}}\n")


def word_block(R: int) -> str:
    """Four steps of the byte-symbol band loop for R rows, one per byte of
    the symbol word, as one block: the compiler puts its own wait state only
    between blocks.  The strand and the conveyor rotate through w, t and u
    (step k moves the strand into the register the conveyor was copied to
    and the conveyor into the strand's old register), and the rows' h
    strands alternate between the caller's registers (h, in place) and a
    scratch set, ending back in h after the fourth step."""
    def cmp(sel, s, a):
        return f'"v_cmp_eq_u32_sdwa %[{s}], %[{a}], %[b] src0_sel:DWORD src1_sel:BYTE_{sel}\\n\\t"'
    regs = [("q", "t", "w"), ("w", "u", "t"), ("t", "w", "u"), ("u", "t", "w")]   # (conveyor, copy/strand, old strand)
    ins = []
    for sel in range(4):
        conv, strand, old = regs[sel]
        src_h, dst_h = ("h", "x") if sel % 2 == 0 else ("x", "h")
        ins += [cmp(sel, "s0", "a0")] + ([cmp(sel, "s1", "a1")] if R > 1 else []) + [
                f'"v_mov_b32_e32 %[{strand}], %[{conv}]\\n\\t"',
                f'"v_mov_b32_dpp %[{strand}], %[{old}] wave_shr:1 row_mask:0xf bank_mask:0xf\\n\\t"',
                f'"v_mov_b32_dpp %[{old}], %[{conv}] wave_shl:1 row_mask:0xf bank_mask:0xf\\n\\t"']
        for k in range(R):
            sk = f"s{k % 2}"
            ins.append(f'"v_min_u32_e32 %[{dst_h}{k}], %[{src_h}{k}], %[{strand}]\\n\\t"')
            ins.append(f'"v_cndmask_b32_e64 %[{dst_h}{k}], %[{dst_h}{k}], %[{strand}], %[{sk}]\\n\\t"')
            ins.append(f'"v_bitop3_b32 %[{strand}], %[{src_h}{k}], %[{strand}], %[{dst_h}{k}] bitop3:0x96\\n\\t"')
            if k + 2 < R:
                ins.append(cmp(sel, sk, f"a{k + 2}"))
    outs = ['[w] "+v"(w)', '[t] "=&v"(t)', '[u] "=&v"(u)'] + [f'[h{k}] "+v"(h[{k}])' for k in range(R)] + \
           [f'[x{k}] "=&v"(x[{k}])' for k in range(R)] + ['[s0] "=&s"(s0)', '[s1] "=&s"(s1)']
    inp = ['[b] "v"(sb)', '[q] "v"(q)'] + [f'[a{k}] "v"(a[{k}])' for k in range(R)]
    body = "\n        ".join(ins)
    return (f"__device__ __forceinline__ void cell_word{R}(const int32_t *a, uint32_t *h, int32_t sb, uint32_t q,"
            f" uint32_t &w, uint32_t &t)\n{{\n    uint64_t s0, s1;\n    uint32_t u, x[{R}];\n    asm volatile(\n"
            f"        {body}\n        : {', '.join(outs)}\n        : {', '.join(inp)});\n}}\n")


def packed_block(R: int) -> str:
    """One step of the packed (two pairs in 16-bit halves) chain for R >= 3
    rows.  Per row k: X e = a ^ b, S mk = sat(1 - e), M hx = sat(mk * 0xffff
    + h), N hn = min(hx, v), B v = h ^ v ^ hn (hn is the row's new h).  A
    packed (VOP3P) result read by the next VALU issue costs a wait state on
    gfx950, so the rows are software-pipelined: row k's N and B run with row
    k+1's M, row k+3's X and row k+2's S between them, in the order
    N(k) M(k+1) B(k) X(k+3) S(k+2), so every packed result is read two or
    more issues later and only the last row's N -> B pair needs an s_nop.
    Four rotating temporaries carry e -> mk -> hx of the rows in flight."""
    assert R >= 3
    def X(k): return f'"v_xor_b32_e32 %[r{k % 4}], %[a{k}], %[b]\\n\\t"'
    def S(k): return f'"v_pk_sub_u16 %[r{k % 4}], 1, %[r{k % 4}] op_sel_hi:[0,1] clamp\\n\\t"'
    def M(k): return f'"v_pk_mad_u16 %[r{k % 4}], %[r{k % 4}], -1, %[h{k}] clamp\\n\\t"'
    def N(k): return f'"v_pk_min_u16 %[n{k}], %[r{k % 4}], %[v]\\n\\t"'
    def B(k): return f'"v_bitop3_b32 %[v], %[h{k}], %[v], %[n{k}] bitop3:0x96\\n\\t"'
    ins = [X(0), S(0), X(1), M(0), S(1), X(2)]
    for k in range(R):
        ins.append(N(k))
        if k + 1 < R:
            ins.append(M(k + 1))
        else:
            ins.append('"s_nop 0\\n\\t"')
        ins.append(B(k))
        if k + 3 < R:
            ins.append(X(k + 3))
        if k + 2 < R:
            ins.append(S(k + 2))
    outs = ['[v] "+v"(v)'] + [f'[n{k}] "=&v"(hn[{k}])' for k in range(R)] + [f'[r{q}] "=&v"(r{q})' for q in range(4)]
    inp = ['[b] "v"(sb)'] + [f'[a{k}] "v"(a[{k}])' for k in range(R)] + [f'[h{k}] "v"(h[{k}])' for k in range(R)]
    body = "\n        ".join(ins)
    return (f"__device__ __forceinline__ void cell16_block{R}(const uint32_t *a, const uint32_t *h, uint32_t *hn,"
            f" uint32_t sb, uint32_t &v)\n{{\n    uint32_t r0, r1, r2, r3;\n    asm volatile(\n"
            f"        {body}\n        : {', '.join(outs)}\n        : {', '.join(inp)});\n}}\n")


def chunk_block(R: int, W: int) -> str:
    """4W steps (W symbol words) of the byte-symbol band loop for R rows as
    one block: the compiler's own wait state before a block that reads the
    strand then comes once per 4W steps.  Step k moves the strand into the
    one of t, u, w that holds neither the strand nor the conveyor (the
    conveyor's copy goes there first), so the roles rotate with period three;
    the rows' h strands alternate between h and a scratch set (4W is even,
    so they end in h).  Returns the strand in `vt` and the conveyor in `qo`.

    With two rows or more the hand-off is two instructions: vcc holds lane 0
    alone, and one `v_cndmask_b32_dpp` (wave_shr:1 with bound_ctrl, so lane
    0 reads 0 and the mask picks the conveyor's lane 0 instead) builds the
    shifted strand; the two compares before it cover the DPP read's two wait
    states.  One row turns the pair round (chunk_block1)."""
    if R == 1:
        return chunk_block1(W)
    def cmp(sel, word, s, a):
        return f'"v_cmp_eq_u32_sdwa %[{s}], %[{a}], %[b{word}] src0_sel:DWORD src1_sel:BYTE_{sel}\\n\\t"'
    ins = ['"s_mov_b64 vcc, 1\\n\\t"'] if R > 1 else []
    strand, conv = "w", "q"           # before step 0: strand in w (the input v), conveyor in q
    for k in range(4 * W):
        word, sel = divmod(k, 4)
        dest = "t" if k == 0 else ({"t", "u", "w"} - {strand, conv}).pop()
        src_h, dst_h = ("h", "x") if k % 2 == 0 else ("x", "h")
        if R > 1:
            hand = [f'"v_cndmask_b32_dpp %[{dest}], %[{strand}], %[{conv}], vcc wave_shr:1 row_mask:0xf'
                    f' bank_mask:0xf bound_ctrl:1\\n\\t"']
        else:
            hand = [f'"v_mov_b32_e32 %[{dest}], %[{conv}]\\n\\t"',
                    f'"v_mov_b32_dpp %[{dest}], %[{strand}] wave_shr:1 row_mask:0xf bank_mask:0xf\\n\\t"']
        ins += [cmp(sel, word, "s0", "a0")] + ([cmp(sel, word, "s1", "a1")] if R > 1 else []) + hand + [
            f'"v_mov_b32_dpp %[{strand}], %[{conv}] wave_shl:1 row_mask:0xf bank_mask:0xf\\n\\t"']
        strand, conv = dest, strand
        for r in range(R):
            sr = f"s{r % 2}"
            ins.append(f'"v_min_u32_e32 %[{dst_h}{r}], %[{src_h}{r}], %[{strand}]\\n\\t"')
            ins.append(f'"v_cndmask_b32_e64 %[{dst_h}{r}], %[{dst_h}{r}], %[{strand}], %[{sr}]\\n\\t"')
            ins.append(f'"v_bitop3_b32 %[{strand}], %[{src_h}{r}], %[{strand}], %[{dst_h}{r}] bitop3:0x96\\n\\t"')
            if r + 2 < R:
                ins.append(cmp(sel, word, sr, f"a{r + 2}"))
    outs = ['[w] "+v"(w)', '[t] "=&v"(t)', '[u] "=&v"(u)'] + [f'[h{r}] "+v"(h[{r}])' for r in range(R)] + \
           [f'[x{r}] "=&v"(x[{r}])' for r in range(R)] + ['[s0] "=&s"(s0)', '[s1] "=&s"(s1)']
    inp = [f'[b{i}] "v"(sb[{i}])' for i in range(W)] + ['[q] "v"(q)'] + [f'[a{r}] "v"(a[{r}])' for r in range(R)]
    body = "\n        ".join(ins)
    clob = ' : "vcc"' if R > 1 else ''
    return (f"__device__ __forceinline__ void cell_chunk{R}x{W}(const int32_t *a, uint32_t *h, const uint32_t *sb,"
            f" uint32_t q, uint32_t &vt, uint32_t &qo)\n{{\n    uint64_t s0, s1;\n    uint32_t w = vt, t, u, x[{R}];\n"
            f"    asm volatile(\n        {body}\n        : {', '.join(outs)}\n        : {', '.join(inp)}"
            f"{clob});\n"
            f"    vt = {strand};\n    qo = {conv};\n}}\n")


def chunk_block1(W: int) -> str:
    """chunk_block for one row.  The single compare cannot cover the two
    wait states between the row's bitop3 and a DPP read of its strand, so
    the conveyor moves first: one `v_cndmask_b32_dpp` shifts it one lane down
    into the free register, vcc holding lane 63 alone so that lane 63 takes
    the strand (a plain operand, which has no wait state); then the compare;
    then the strand moves one lane up in place of the old conveyor, whose
    lane 0 (the band input) the DPP keeps.  Three VALU per step besides the
    cell's three; strand, conveyor and free register rotate through w, q and
    t with period three."""
    def cmp(sel, word):
        return f'"v_cmp_eq_u32_sdwa %[s0], %[a0], %[b{word}] src0_sel:DWORD src1_sel:BYTE_{sel}\\n\\t"'
    ins = ['"s_mov_b32 vcc_lo, 0\\n\\t"', '"s_mov_b32 vcc_hi, 0x80000000\\n\\t"']
    strand, conv, free = "w", "q", "t"
    for k in range(4 * W):
        word, sel = divmod(k, 4)
        src_h, dst_h = ("h", "x") if k % 2 == 0 else ("x", "h")
        move_conv = (f'"v_cndmask_b32_dpp %[{free}], %[{conv}], %[{strand}], vcc wave_shl:1 row_mask:0xf'
                     f' bank_mask:0xf bound_ctrl:1\\n\\t"')
        move_strand = f'"v_mov_b32_dpp %[{conv}], %[{strand}] wave_shr:1 row_mask:0xf bank_mask:0xf\\n\\t"'
        ins += ([cmp(sel, word), move_conv] if k == 0 else [move_conv, cmp(sel, word)]) + [move_strand]
        strand, conv, free = conv, free, strand
        ins.append(f'"v_min_u32_e32 %[{dst_h}0], %[{src_h}0], %[{strand}]\\n\\t"')
        ins.append(f'"v_cndmask_b32_e64 %[{dst_h}0], %[{dst_h}0], %[{strand}], %[s0]\\n\\t"')
        ins.append(f'"v_bitop3_b32 %[{strand}], %[{src_h}0], %[{strand}], %[{dst_h}0] bitop3:0x96\\n\\t"')
    outs = ['[w] "+v"(w)', '[q] "+v"(q)', '[t] "=&v"(t)', '[h0] "+v"(h[0])', '[x0] "=&v"(x0)', '[s0] "=&s"(s0)']
    inp = [f'[b{i}] "v"(sb[{i}])' for i in range(W)] + ['[a0] "v"(a[0])']
    body = "\n        ".join(ins)
    return (f"__device__ __forceinline__ void cell_chunk1x{W}(const int32_t *a, uint32_t *h, const uint32_t *sb,"
            f" uint32_t q, uint32_t &vt, uint32_t &qo)\n{{\n    uint64_t s0;\n    uint32_t w = vt, t, x0;\n"
            f"    asm volatile(\n        {body}\n        : {', '.join(outs)}\n        : {', '.join(inp)} : \"vcc\");\n"
            f"    vt = {strand};\n    qo = {conv};\n}}\n")


def packed_chunk(R: int, steps: int = 32, oh: bool = False) -> str:
    """`steps` steps of the packed chain for R >= 3 rows as one block, the
    hand-off included.  Per step: the symbols (bq one lane up into IB's
    register, lane 0 keeping IB's lane 0; the next step's IB, IB one lane
    down with lane 63 taking 0, was made during the previous step), then the
    strand (one `v_cndmask_b32_dpp` with wave_shr:1 and bound_ctrl moves v
    one lane up into the free register, vcc holding lane 0 alone so that
    lane 0 takes the conveyor's lane 0; the conveyor one lane down in place
    of v, lane 63 keeping v), then the rows as in packed_block.  The next
    step's IB move fills the last row's min -> bitop3 slot (an s_nop in the
    chunk's last step, whose IB the caller may reload), and the symbol move
    and the first xor put two issues between the last bitop3 and the DPP
    that reads its strand, so no other step needs an s_nop.  Strand,
    conveyor and the free register rotate through three registers, IB, bq
    and the free one through three others, both with period three.

    oh: one-hot symbols (each 16-bit half holds 1 << symbol, padding 0, for
    alphabets of at most 16).  A row's match mask is then a & b, nonzero on
    a match, and sat(mk * 0xffff + h) is 0xffff there: four VALU per row
    (v_and, v_pk_mad_u16 clamp, v_pk_min_u16, v_bitop3) in the order
    N(k) M(k+1) B(k) A(k+2), every packed result again read two issues
    later."""
    assert R >= 3
    S, K, F = "va", "vb", "vc"        # strand, conveyor, free
    I, B, G = "ia", "ib", "ic"        # symbol conveyor IB, symbols bq, free
    op = "v_and_b32_e32" if oh else "v_xor_b32_e32"
    def X(k, sym): return f'"{op} %[r{k % 4}], %[a{k}], %[{sym}]\\n\\t"'
    def Sb(k): return f'"v_pk_sub_u16 %[r{k % 4}], 1, %[r{k % 4}] op_sel_hi:[0,1] clamp\\n\\t"'
    def M(k, hs): return f'"v_pk_mad_u16 %[r{k % 4}], %[r{k % 4}], -1, %[{hs}{k}] clamp\\n\\t"'
    def N(k, hd, st): return f'"v_pk_min_u16 %[{hd}{k}], %[r{k % 4}], %[{st}]\\n\\t"'
    def Bx(k, hs, hd, st): return f'"v_bitop3_b32 %[{st}], %[{hs}{k}], %[{st}], %[{hd}{k}] bitop3:0x96\\n\\t"'
    def ib_next(dst, src): return f'"v_mov_b32_dpp %[{dst}], %[{src}] wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\\n\\t"'
    ins = ['"s_mov_b64 vcc, 1\\n\\t"', '"s_nop 1\\n\\t"', ib_next(G, I)]
    for st in range(steps):
        hs, hd = ("h", "x") if st % 2 == 0 else ("x", "h")
        # symbols bq move into I (lane 0 keeps IB's lane 0); the next IB is already in G
        ins += [f'"v_mov_b32_dpp %[{I}], %[{B}] wave_shr:1 row_mask:0xf bank_mask:0xf\\n\\t"']
        sym = I
        ins += [X(0, sym),
                f'"v_cndmask_b32_dpp %[{F}], %[{S}], %[{K}], vcc wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\\n\\t"',
                f'"v_mov_b32_dpp %[{S}], %[{K}] wave_shl:1 row_mask:0xf bank_mask:0xf\\n\\t"']
        strand, conv = F, S
        if oh:
            ins += [M(0, hs), X(1, sym)]
        else:
            ins += [Sb(0), X(1, sym), M(0, hs), Sb(1), X(2, sym)]
        for k in range(R):
            ins.append(N(k, hd, strand))
            if k + 1 < R:
                ins.append(M(k + 1, hs))
            elif st + 1 < steps:
                ins.append(ib_next(B, G))   # the next step's IB (G one lane down) into the free register
            else:
                ins.append('"s_nop 0\\n\\t"')
            ins.append(Bx(k, hs, hd, strand))
            if oh:
                if k + 2 < R:
                    ins.append(X(k + 2, sym))
                continue
            if k + 3 < R:
                ins.append(X(k + 3, sym))
            if k + 2 < R:
                ins.append(Sb(k + 2))
        S, K, F = strand, conv, K
        I, B, G = G, I, B
    regs = {"va": "v", "vb": "q", "vc": "c", "ia": "ib", "ib": "bq", "ic": "g"}
    outs = ['[va] "+v"(v)', '[vb] "+v"(q)', '[vc] "=&v"(c)', '[ia] "+v"(ib)', '[ib] "+v"(bq)', '[ic] "=&v"(g)'] + \
           [f'[h{k}] "+v"(h[{k}])' for k in range(R)] + [f'[x{k}] "=&v"(x[{k}])' for k in range(R)] + \
           [f'[r{q}] "=&v"(r{q})' for q in range(4)]
    inp = [f'[a{k}] "v"(a[{k}])' for k in range(R)]
    body = "\n        ".join(ins)
    name = f"cell16oh_chunk{R}" if oh else f"cell16_chunk{R}"
    return (f"__device__ __forceinline__ void {name}(const uint32_t *a, uint32_t *h, uint32_t &sv,"
            f" uint32_t &sq, int32_t &sib, int32_t &sbq)\n{{\n"
            f"    uint32_t v = sv, q = sq, c, ib = (uint32_t)sib, bq = (uint32_t)sbq, g, r0, r1, r2, r3, x[{R}];\n"
            f"    asm volatile(\n        {body}\n        : {', '.join(outs)}\n        : {', '.join(inp)} : \"vcc\");\n"
            f"    sv = {regs[S]};\n    sq = {regs[K]};\n    sib = (int32_t){regs[I]};\n    sbq = (int32_t){regs[B]};\n}}\n")


if __name__ == "__main__":
    print("// Generated by gen_cell_asm.py -- hand-scheduled cell blocks (see there).")
    print("#pragma once\n#include <stdint.h>\n")
    for R in (4, 8, 16):
        print(block(R))
    for R in (4, 8, 16):
        for q in range(4):
            print(block(R, q))
    print("// cell_block{R}_byte{SEL} by template arguments")
    print("template <int R, int SEL>\n__device__ __forceinline__ void cell_block_byte(const int32_t *a, const uint32_t *h,"
          " uint32_t *hn, int32_t sb, uint32_t &v)\n{")
    first = True
    for R in (4, 8, 16):
        for q in range(4):
            kw = "if constexpr" if first else "else if constexpr"
            print(f"    {kw} (R == {R} && SEL == {q}) cell_block{R}_byte{q}(a, h, hn, sb, v);")
            first = False
    print('    else static_assert(R == 4 || R == 8 || R == 16, "rows per block");\n}\n')
    for R in (4, 8, 16):
        for q in range(4):
            print(step_block(R, q))
    for R in (1, 2, 4, 8, 16):
        print(word_block(R))
    for R in (4, 8, 16):
        print(packed_block(R))
    for R in (4, 8, 16):
        print(packed_chunk(R))
    print("// cell16_chunk{R} by template argument: the 32 steps of a packed chunk")
    print("template <int R>\n__device__ __forceinline__ void cell16_chunk(const uint32_t *a, uint32_t *h, uint32_t &v,"
          " uint32_t &q, int32_t &ib, int32_t &bq)\n{")
    print("    if constexpr (R == 4) cell16_chunk4(a, h, v, q, ib, bq);")
    print("    else if constexpr (R == 8) cell16_chunk8(a, h, v, q, ib, bq);")
    print("    else if constexpr (R == 16) cell16_chunk16(a, h, v, q, ib, bq);")
    print('    else static_assert(R == 4 || R == 8 || R == 16, "rows per block");\n}\n')
    for R in (4, 8, 16):
        print(packed_chunk(R, oh=True))
    print("// cell16oh_chunk{R} by template argument: the 32 steps of a packed chunk, one-hot symbols")
    print("template <int R>\n__device__ __forceinline__ void cell16oh_chunk(const uint32_t *a, uint32_t *h, uint32_t &v,"
          " uint32_t &q, int32_t &ib, int32_t &bq)\n{")
    print("    if constexpr (R == 4) cell16oh_chunk4(a, h, v, q, ib, bq);")
    print("    else if constexpr (R == 8) cell16oh_chunk8(a, h, v, q, ib, bq);")
    print("    else if constexpr (R == 16) cell16oh_chunk16(a, h, v, q, ib, bq);")
    print('    else static_assert(R == 4 || R == 8 || R == 16, "rows per block");\n}\n')
    print("// cell16_block{R} by template argument")
    print("template <int R>\n__device__ __forceinline__ void cell16_block(const uint32_t *a, const uint32_t *h,"
          " uint32_t *hn, uint32_t sb, uint32_t &v)\n{")
    print("    if constexpr (R == 4) cell16_block4(a, h, hn, sb, v);")
    print("    else if constexpr (R == 8) cell16_block8(a, h, hn, sb, v);")
    print("    else if constexpr (R == 16) cell16_block16(a, h, hn, sb, v);")
    print('    else static_assert(R == 4 || R == 8 || R == 16, "rows per block");\n}\n')
    print("// cell_word{R} by template argument")
    print("template <int R>\n__device__ __forceinline__ void cell_word(const int32_t *a, uint32_t *h, int32_t sb,"
          " uint32_t q, uint32_t &w, uint32_t &t)\n{")
    print("    if constexpr (R == 1) cell_word1(a, h, sb, q, w, t);")
    print("    else if constexpr (R == 2) cell_word2(a, h, sb, q, w, t);")
    print("    else if constexpr (R == 4) cell_word4(a, h, sb, q, w, t);")
    print("    else if constexpr (R == 8) cell_word8(a, h, sb, q, w, t);")
    print("    else if constexpr (R == 16) cell_word16(a, h, sb, q, w, t);")
    print('    else static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 16, "rows per block");\n}\n')
    for R in (1, 2, 4, 8, 16):
        print(chunk_block(R, 8))
    print("// cell_chunk{R}x8 by template argument: the 32 steps of a chunk")
    print("template <int R>\n__device__ __forceinline__ void cell_chunk32(const int32_t *a, uint32_t *h,"
          " const uint32_t *sb, uint32_t q, uint32_t &vt, uint32_t &qo)\n{")
    print("    if constexpr (R == 1) cell_chunk1x8(a, h, sb, q, vt, qo);")
    for R in (2, 4, 8, 16):
        print(f"    else if constexpr (R == {R}) cell_chunk{R}x8(a, h, sb, q, vt, qo);")
    print('    else static_assert(R == 1 || R == 2 || R == 4 || R == 8 || R == 16, "rows per block");\n}\n')
    print("// cell_step{R}_byte{SEL} by template arguments")
    print("template <int R, int SEL>\n__device__ __forceinline__ void cell_step_byte(const int32_t *a, const uint32_t *h,"
          " uint32_t *hn, int32_t sb, uint32_t q, uint32_t &w, uint32_t &t)\n{")
    first = True
    for R in (4, 8, 16):
        for q in range(4):
            kw = "if constexpr" if first else "else if constexpr"
            print(f"    {kw} (R == {R} && SEL == {q}) cell_step{R}_byte{q}(a, h, hn, sb, q, w, t);")
            first = False
    print('    else static_assert(R == 4 || R == 8 || R == 16, "rows per block");\n}')
