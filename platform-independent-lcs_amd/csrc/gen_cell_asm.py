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
        ins += [cmp(sel, "s0", "a0"), cmp(sel, "s1", "a1"),
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
    for R in (4, 8, 16):
        print(word_block(R))
    print("// cell_word{R} by template argument")
    print("template <int R>\n__device__ __forceinline__ void cell_word(const int32_t *a, uint32_t *h, int32_t sb,"
          " uint32_t q, uint32_t &w, uint32_t &t)\n{")
    print("    if constexpr (R == 4) cell_word4(a, h, sb, q, w, t);")
    print("    else if constexpr (R == 8) cell_word8(a, h, sb, q, w, t);")
    print("    else if constexpr (R == 16) cell_word16(a, h, sb, q, w, t);")
    print('    else static_assert(R == 4 || R == 8 || R == 16, "rows per block");\n}\n')
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
