// Seaweed (strand) combing kernels for CDNA4 / gfx950.
//
// Grid orientation used on the device: row r = 0..m-1 holds a[r] (the
// reference's reversed row i = m-1-r, src/lcs_reference.hpp:13-20), column c
// holds b[c].  Horizontal strands travel left to right along a row, vertical
// strands travel top to bottom (from row r-1 into row r, the reference's "from
// high i to low i", src/lcs_common.hpp:76-88).  Cell rule
// (src/lcs_common.hpp:7-24): swap iff a[r] == b[c] || h > v.  Strand ids are
// non-negative, so the signed compare equals the unsigned one and the swap is
//     h' = min(match ? ~0u : h, v),   v' = h ^ v ^ h'
// i.e. v_cmp_eq + v_min_u32 + v_cndmask + v_bitop3 (xor3): 4 VALU, no SALU.
//
// Band chain.  One wave64 owns a band of 64*TM consecutive rows; lane L keeps
// TM rows (symbols and h strands) in VGPRs.  At step t lane L combs column
// t-L, so the wave sweeps the band along anti-diagonals.  The v strand and the
// b symbol move one lane down per step by DPP wave_shr:1; lane 0 receives the
// band's inputs from two "conveyor" registers shifted by wave_shl:1, and the
// strand conveyor also catches the v strand leaving lane 63: a chunk's C
// inputs drain out at lane 0 while its C outputs fill lanes 64-C .. 63.
// Nothing touches LDS.
// Every C = 32 steps (a chunk) the wave publishes the C newest bottom-edge
// strands of its band to a ring in global memory as 64-bit (tag, strand) words stored with
// device-scope relaxed atomics; the band below polls for words carrying its
// tag, so no fence, flag or cache flush is needed.  Bands are handed out by an
// atomic ticket in launch order, so a wave only ever waits for a band that a
// resident wave already owns: any grid size makes progress.  Every wait is
// bounded in time; on expiry the kernel raises an abort flag that every wave
// checks, and all waves exit.
//
// Residuals (the reference's CPU fixup, src/lcs_grid.hpp:451-493) are handled
// in the kernel: lanes whose column is outside [0, n) skip the cells, rows
// beyond m in the last band never match and hold the smallest strand id, so
// they pass every v strand through unchanged.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "cell_asm.h"
#include "comb_kernels.h"

namespace lcsk {

constexpr int kWave = 64;
constexpr int kDppWaveShl1 = 0x130;  // lane L reads lane L+1
constexpr int kDppWaveShr1 = 0x138;  // lane L reads lane L-1

__device__ __forceinline__ int32_t shl1_zero_i(int32_t x)
{
    return __builtin_amdgcn_mov_dpp(x, kDppWaveShl1, 0xf, 0xf, true);
}
// lane L <- x[L-1]; lane 0 keeps `old`
__device__ __forceinline__ uint32_t shr1_keep(uint32_t old_, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old_, (int)x, kDppWaveShr1, 0xf, 0xf, false);
}
__device__ __forceinline__ int32_t shr1_keep_i(int32_t old_, int32_t x)
{
    return __builtin_amdgcn_update_dpp(old_, x, kDppWaveShr1, 0xf, 0xf, false);
}
// lane L <- x[L+1]; lane 63 keeps `old`
__device__ __forceinline__ uint32_t shl1_keep(uint32_t old_, uint32_t x)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old_, (int)x, kDppWaveShl1, 0xf, 0xf, false);
}

// Pointers read from a pair descriptor are generic; the band loop casts them
// to the global address space so its loads and stores are global_* rather
// than flat_* instructions.  A flat access may hit LDS, whose results return
// out of order, and that forces a full vmcnt(0) wait (on the in-flight ring
// store too) where a global access needs only vmcnt(1).
#define LCSK_GLOBAL __attribute__((address_space(1)))
typedef LCSK_GLOBAL unsigned long long gword;
typedef LCSK_GLOBAL const unsigned long long gcword;
typedef LCSK_GLOBAL const int32_t gsym;
typedef LCSK_GLOBAL const uint32_t gsym8;   // a word of four byte symbols
typedef LCSK_GLOBAL int32_t gout;
// LDS words polled through a pointer: the explicit address space keeps the
// reads ds_read (a volatile generic pointer would be a flat access, and a
// flat access waits for every outstanding global store).
typedef __attribute__((address_space(3))) volatile unsigned lds_flag;

__device__ __forceinline__ unsigned long long ld_relaxed(gcword *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(gword *p, unsigned long long x)
{
    __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_flag(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void raise_abort(unsigned *p)
{
    __hip_atomic_store(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One cell.  `sb` is the column symbol; `valid` is false for padding rows.
template <bool ROWMASK>
__device__ __forceinline__ void cell(int32_t sa, int32_t sb, bool valid, uint32_t &h, uint32_t &v)
{
    bool match = (sa == sb);
    if (ROWMASK) match = match && valid;
    uint32_t hx = match ? 0xffffffffu : h;
    uint32_t hn = min(hx, v);
    v = __builtin_amdgcn_bitop3_b32(h, v, hn, 0x96);
    h = hn;
}

// All TM rows of one column for one lane.
template <int TM, bool ROWMASK>
__device__ __forceinline__ void column_cells(const int32_t (&a)[TM], uint32_t (&h)[TM], const bool (&valid)[TM],
                                             int32_t sb, uint32_t &v)
{
    // Rows go through hand-scheduled blocks of up to 16 (cell_asm.h); one
    // block per column keeps the whole chain free of hazard nops.
    if constexpr (!ROWMASK && TM >= 4) {
        constexpr int R = TM >= 16 ? 16 : TM;
#pragma unroll
        for (int k = 0; k < TM; k += R) {
            uint32_t hn[R];
            if constexpr (R == 16) cell_block16(&a[k], &h[k], hn, sb, v);
            else if constexpr (R == 8) cell_block8(&a[k], &h[k], hn, sb, v);
            else cell_block4(&a[k], &h[k], hn, sb, v);
#pragma unroll
            for (int q = 0; q < R; ++q) h[k + q] = hn[q];
        }
    } else {
#pragma unroll
        for (int k = 0; k < TM; ++k) cell<ROWMASK>(a[k], sb, valid[k], h[k], v);
    }
}

// Two cells at once, pair X in the low and pair Y in the high 16 bits of
// every register (the packed chain): per half, swap iff sa == sb || h > v,
// as the 32-bit cell.  With e = sa ^ sb, sat(1 - e) is 1 on a match, and
// sat(h + 1 * 0xffff) = 0xffff replaces h there, so
//     hx = mad_sat(sat_sub(1, e), 0xffff, h),  hn = min(hx, v),  v' = h ^ v ^ hn
// is v_xor + v_pk_sub_u16 clamp + v_pk_mad_u16 clamp + v_pk_min_u16 +
// v_bitop3: 5 VALU for 2 cells (4 per cell in 32 bits).  Strand ids stay
// below 0xffff, the value a padded column carries.
typedef unsigned short lcsk_u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void cell16(uint32_t sa, uint32_t sb, uint32_t &h, uint32_t &v)
{
    const lcsk_u16x2 one = {1, 1};
    const lcsk_u16x2 mk = __builtin_elementwise_sub_sat(one, __builtin_bit_cast(lcsk_u16x2, sa ^ sb));
    uint32_t hx;
    asm("v_pk_mad_u16 %0, %1, -1, %2 clamp" : "=v"(hx) : "v"(__builtin_bit_cast(uint32_t, mk)), "v"(h));
    const uint32_t hn = __builtin_bit_cast(
        uint32_t, __builtin_elementwise_min(__builtin_bit_cast(lcsk_u16x2, hx), __builtin_bit_cast(lcsk_u16x2, v)));
    v = __builtin_amdgcn_bitop3_b32(h, v, hn, 0x96);
    h = hn;
}

// The packed cell with one-hot symbols (each half holds 1 << symbol, 0 for
// padding; alphabets of at most 16): sa & sb is nonzero exactly on a match,
// and sat(mk * 0xffff + h) is 0xffff there, so the match mask costs one
// v_and instead of v_xor + v_pk_sub_u16: 4 VALU for 2 cells.
__device__ __forceinline__ void cell16_oh(uint32_t sa, uint32_t sb, uint32_t &h, uint32_t &v)
{
    uint32_t hx;
    asm("v_pk_mad_u16 %0, %1, -1, %2 clamp" : "=v"(hx) : "v"(sa & sb), "v"(h));
    const uint32_t hn = __builtin_bit_cast(
        uint32_t, __builtin_elementwise_min(__builtin_bit_cast(lcsk_u16x2, hx), __builtin_bit_cast(lcsk_u16x2, v)));
    v = __builtin_amdgcn_bitop3_b32(h, v, hn, 0x96);
    h = hn;
}

// Steps t0 .. t0+C-1 of a band.  COLMASK: some lane's column leaves [0, n).
// Q (strands) and IB (symbols) are the conveyors.  On entry Q's lanes
// 0 .. C-1 hold the chunk's inputs (the input of step t0+s in lane s); each
// step hands Q's lane 0 to the band and shifts Q down one lane while lane
// 63's strand enters at the top, so on return Q's lanes 64-C .. 63 hold the
// chunk's outputs (columns t0-64 .. t0-64+C-1).  One shift serves both
// directions: C <= 64, so the inputs and outputs never meet.
template <int TM, int C, bool ROWMASK, bool COLMASK, bool PK = false, bool OH = false>
__device__ __forceinline__ void band_steps(const int32_t (&a)[TM], uint32_t (&h)[TM], const bool (&valid)[TM],
                                           uint32_t &v, int32_t &bq, uint32_t &Q, int32_t &IB, int t0, int lane, int n)
{
#ifndef LCS_PACKED_CHUNK_BLOCKS
#define LCS_PACKED_CHUNK_BLOCKS 1
#endif
    if constexpr (LCS_PACKED_CHUNK_BLOCKS && PK && !COLMASK && C == 32 && TM >= 4 && TM <= 16) {
        // the packed chain's 32 steps, hand-offs included, in one asm block
        // (cell_asm.h): no step waits out a hazard with an s_nop
        uint32_t ua[TM];
#pragma unroll
        for (int q = 0; q < TM; ++q) ua[q] = (uint32_t)a[q];
        if constexpr (OH) cell16oh_chunk<TM>(ua, h, v, Q, IB, bq);
        else cell16_chunk<TM>(ua, h, v, Q, IB, bq);
        return;
    }
#pragma unroll
    for (int s = 0; s < C; ++s) {
        const int32_t bin = IB;
        IB = shl1_zero_i(IB);
        const uint32_t vn = shr1_keep(Q, v);   // lane 0 takes the band input Q[0]
        Q = shl1_keep(v, Q);                   // inputs move toward lane 0; lane 63's output enters at the top
        v = vn;
        bq = shr1_keep_i(bin, bq);
        bool active = true;
        if (COLMASK) {
            const int j = t0 + s - lane;
            active = (j >= 0) && (j < n);
        }
        if (active) {
            if constexpr (PK && OH) {
#pragma unroll
                for (int k = 0; k < TM; ++k) cell16_oh((uint32_t)a[k], (uint32_t)bq, h[k], v);
            } else if constexpr (PK && !COLMASK && TM >= 4) {
                // software-pipelined rows (cell_asm.h): no packed result is
                // read by the next issue, so no s_nop between the rows
                constexpr int R = TM >= 16 ? 16 : TM;
#pragma unroll
                for (int k = 0; k < TM; k += R) {
                    uint32_t ua[R], hn[R];
#pragma unroll
                    for (int q = 0; q < R; ++q) ua[q] = (uint32_t)a[k + q];
                    cell16_block<R>(ua, &h[k], hn, (uint32_t)bq, v);
#pragma unroll
                    for (int q = 0; q < R; ++q) h[k + q] = hn[q];
                }
            } else if constexpr (PK) {
#pragma unroll
                for (int k = 0; k < TM; ++k) cell16((uint32_t)a[k], (uint32_t)bq, h[k], v);
            } else {
                column_cells<TM, ROWMASK>(a, h, valid, bq, v);
            }
        }
    }
}

// All TM rows of one column for one lane when the column symbol is byte SEL
// of `word` (byte symbols; a row symbol matches it zero-extended, so row
// symbols outside [0, 255] never match, as no column symbol is).
template <int TM, bool ROWMASK, int SEL>
__device__ __forceinline__ void column_cells8(const int32_t (&a)[TM], uint32_t (&h)[TM], const bool (&valid)[TM],
                                              uint32_t word, uint32_t &v)
{
    if constexpr (!ROWMASK && TM >= 4) {
        constexpr int R = TM >= 16 ? 16 : TM;
#pragma unroll
        for (int k = 0; k < TM; k += R) {
            uint32_t hn[R];
            cell_block_byte<R, SEL>(&a[k], &h[k], hn, (int32_t)word, v);
#pragma unroll
            for (int q = 0; q < R; ++q) h[k + q] = hn[q];
        }
    } else {
        const int32_t sb = (int32_t)((word >> (8 * SEL)) & 0xffu);
#pragma unroll
        for (int k = 0; k < TM; ++k) cell<ROWMASK>(a[k], sb, valid[k], h[k], v);
    }
}

// Step t of a band with byte symbols: the strand moves as in band_steps;
// the column symbol is byte SEL of `word`, which the lane holds itself.
// An unmasked step of four or more rows runs as one asm block (cell_asm.h:
// the strand hand-off sits behind the step's first two compares, so the DPP
// read of v needs no s_nop after the previous step's last cell wrote it);
// band_steps8 uses this for 32-row tiles, and for every tile height when
// its four-step blocks are compiled out.
template <int TM, bool ROWMASK, bool COLMASK, int SEL>
__device__ __forceinline__ void band_step8(const int32_t (&a)[TM], uint32_t (&h)[TM], const bool (&valid)[TM],
                                           uint32_t &v, uint32_t &Q, uint32_t word, int t, int lane, int n)
{
    if constexpr (!ROWMASK && !COLMASK && TM >= 4) {
        constexpr int R = TM >= 16 ? 16 : TM;
        uint32_t w = v, vt;
        {
            uint32_t hn[R];
            cell_step_byte<R, SEL>(&a[0], &h[0], hn, (int32_t)word, Q, w, vt);
#pragma unroll
            for (int q = 0; q < R; ++q) h[q] = hn[q];
        }
#pragma unroll
        for (int k = R; k < TM; k += R) {
            uint32_t hn[R];
            cell_block_byte<R, SEL>(&a[k], &h[k], hn, (int32_t)word, vt);
#pragma unroll
            for (int q = 0; q < R; ++q) h[k + q] = hn[q];
        }
        Q = w;    // moved down one lane, lane 63 holding the strand that left the band
        v = vt;
        return;
    }
    const uint32_t vn = shr1_keep(Q, v);
    Q = shl1_keep(v, Q);
    v = vn;
    bool active = true;
    if (COLMASK) {
        const int j = t - lane;
        active = (j >= 0) && (j < n);
    }
    if (active) column_cells8<TM, ROWMASK, SEL>(a, h, valid, word, v);
}

// Steps t0 .. t0+C-1 of a band whose column symbols are bytes: sym[s / 4]
// byte s % 4 is the symbol of the column this lane combs at step t0 + s
// (every lane holds its own C symbols), so only the strand conveyor Q
// moves between lanes.
template <int TM, int C, bool ROWMASK, bool COLMASK>
__device__ __forceinline__ void band_steps8(const int32_t (&a)[TM], uint32_t (&h)[TM], const bool (&valid)[TM],
                                            uint32_t &v, uint32_t &Q, const uint32_t (&sym)[C / 4], int t0, int lane,
                                            int n)
{
// The compiler puts an s_nop 0 before every asm block that reads the strand,
// so larger blocks pay it less often.  LCS_CHUNK_BLOCKS=0 falls back to one
// block per symbol word and LCS_WORD_BLOCKS=0 then to one per step (m = n =
// 2^20: 131.5 ms per comb with chunk blocks, 132.0 with word blocks on the
// same box; 133.6 with step blocks against 132.1 for word blocks on another).
#ifndef LCS_WORD_BLOCKS
#define LCS_WORD_BLOCKS 1
#endif
#ifndef LCS_CHUNK_BLOCKS
#define LCS_CHUNK_BLOCKS 1
#endif
    if constexpr (LCS_CHUNK_BLOCKS && C == 32 && !ROWMASK && !COLMASK && TM <= 16) {
        // the whole chunk in one asm block (cell_asm.h)
        uint32_t qn;
        cell_chunk32<TM>(a, h, sym, Q, v, qn);
        Q = qn;
        return;
    }
    if constexpr (LCS_WORD_BLOCKS && !ROWMASK && !COLMASK && TM <= 16) {
        // the four steps of a symbol word in one asm block (cell_asm.h)
#pragma unroll
        for (int w = 0; w < C / 4; ++w) {
            uint32_t vw = v, vt;
            cell_word<TM>(a, h, (int32_t)sym[w], Q, vw, vt);
            Q = vw;
            v = vt;
        }
        return;
    }
#pragma unroll
    for (int w = 0; w < C / 4; ++w) {
        const int t = t0 + 4 * w;
        band_step8<TM, ROWMASK, COLMASK, 0>(a, h, valid, v, Q, sym[w], t, lane, n);
        band_step8<TM, ROWMASK, COLMASK, 1>(a, h, valid, v, Q, sym[w], t + 1, lane, n);
        band_step8<TM, ROWMASK, COLMASK, 2>(a, h, valid, v, Q, sym[w], t + 2, lane, n);
        band_step8<TM, ROWMASK, COLMASK, 3>(a, h, valid, v, Q, sym[w], t + 3, lane, n);
    }
}

// Every band runs the same sequence of memory operations: it reads its top
// edge from ring slot g&1 (tag g) and writes its bottom edge to slot (g+1)&1
// (tag g+1).  The prep kernel seeds slot 0 with the grid's top edge (tag 0)
// and the extract kernel reads v_out from the last band's slot, so no
// branch on the band's position sits between a load and its wait.
// PK: the packed chain over a CoupleDesc -- every symbol and strand
// register holds pair X in its low and pair Y in its high 16 bits.  Rows
// past a pair's m hold symbol kPadRow16 and strand 0 and columns past its n
// carry symbol kPadCol16 and strand 0xffff: they never match and never
// swap, so they pass every strand through unchanged and ROWMASK is not
// needed.
constexpr uint32_t kPadRow16 = 0xffffu, kPadCol16 = 0xfffeu;

// S8: the column symbols come from the byte copies of b (P.sym8) instead of
// b; every lane loads its own C symbols of the next chunk (C/4 words from
// the copy that starts its window on a word boundary) and no symbol
// conveyor runs.
template <int TM, int C, bool ROWMASK, bool PK, typename Desc, bool S8 = false, bool OH = false>
__device__ __forceinline__ bool band_pass(const ChainArgs &A, const Desc &P, int g, int lane, int t_id)
{
    const int m = P.m, n = P.n;
    const int r0 = g * (kWave * TM) + lane * TM;
    if (A.trace && t_id < A.trace_cap && lane == 0) A.trace[2 * (size_t)t_id] = __builtin_amdgcn_s_memrealtime();
    int32_t a[TM];
    uint32_t h[TM];
    bool valid[TM];
#pragma unroll
    for (int k = 0; k < TM; ++k) {
        const int r = r0 + k;
        if constexpr (PK) {
            uint32_t sa = 0, hh = 0;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const bool in = r < P.m_[q];
                uint32_t sym;
                if constexpr (OH) sym = in ? 1u << ((uint32_t)P.a[q][r] & 15u) : 0u;
                else sym = in ? ((uint32_t)P.a[q][r] & 0xffffu) : kPadRow16;
                const uint32_t id = in ? (uint32_t)(P.m_[q] - 1 - r) : 0u;
                sa |= sym << (16 * q);
                hh |= id << (16 * q);
            }
            valid[k] = true;
            a[k] = (int32_t)sa;
            h[k] = hh;
        } else if constexpr (S8) {
            // rows past m hold a symbol no byte equals and strand 0: they
            // never match and never swap (min(0, v) = 0), so every strand
            // passes them unchanged and the band needs no row mask
            valid[k] = r < m;
            a[k] = valid[k] ? P.a[r] : -1;
            h[k] = valid[k] ? (P.h_in ? (uint32_t)P.h_in[m - 1 - r] : (uint32_t)(m - 1 - r)) : 0u;
        } else {
            valid[k] = ROWMASK ? (r < m) : true;
            a[k] = valid[k] ? P.a[r] : 0;
            h[k] = valid[k] ? (P.h_in ? (uint32_t)P.h_in[m - 1 - r] : (uint32_t)(m - 1 - r)) : 0u;
        }
    }
    const uint32_t tag_in = (uint32_t)g;
    const unsigned long long tag_out = (unsigned long long)(g + 1) << 32;
    gcword *Ein = (gcword *)(P.E + (size_t)(g & 1) * (size_t)n);
    gword *Eout = (gword *)(P.E + (size_t)((g + 1) & 1) * (size_t)n);
    const int nchunks = (n + kWave + C - 1) / C;
    // the column symbol of column c (c in [0, n)): for the packed chain both
    // pairs' symbols, each from a clamped in-bounds address
    auto load_b = [&](auto c) -> int32_t {
        if constexpr (PK) {
            uint32_t x = 0;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const uint32_t raw = (uint32_t)((gsym *)P.b[q])[min((int)c, P.n_[q] - 1)];
                if constexpr (OH) x |= ((int)c < P.n_[q] ? 1u << (raw & 15u) : 0u) << (16 * q);
                else x |= ((int)c < P.n_[q] ? raw & 0xffffu : kPadCol16) << (16 * q);
            }
            return (int32_t)x;
        } else {
            return ((gsym *)P.b)[c];
        }
    };

    // Byte symbols: lane L combs column t - L at step t, the byte at
    // x0 + t of the padded symbol string (x0 = kSym8Pad + col0 - L >= 65);
    // copy j = x0 % 4 holds it at byte x0 - j + t, a word boundary for
    // every chunk (t0 is a multiple of C).  The copies extend kSym8Pad
    // bytes past column n, beyond the last prefetch (t0 <= n + 64 + C).
    static_assert(!S8 || (C % 4 == 0 && kWave + 2 * C - 2 < kSym8Pad), "byte-symbol window");
    [[maybe_unused]] uint32_t sym[C / 4], psym[C / 4];
    [[maybe_unused]] gsym8 *sym_base = nullptr;
    if constexpr (S8) {
        const int x0 = kSym8Pad + P.sym8_col0 - lane;
        const int j = x0 & 3;
        sym_base = (gsym8 *)P.sym8 + (size_t)j * (size_t)P.sym8_stride + ((x0 - j) >> 2);
    }
    auto load_sym = [&](uint32_t (&d)[C / 4], int t) {
        gsym8 *p = sym_base + (t >> 2);
#pragma unroll
        for (int q = 0; q < C / 4; ++q) d[q] = p[q];
    };
    // prefetch chunk 0.  Every lane loads from a clamped in-bounds address
    // (no exec masking, no select), so the waits before each chunk cover
    // exactly these loads.
    const int nm1 = n - 1;
    // Only lanes < C use their words (the conveyors start at lane 0 and
    // shift C times); lanes >= C load the same addresses, so each chunk
    // fetches C words, not 64.
    int32_t pB = 0;
    if constexpr (S8) load_sym(sym, 0);
    else pB = load_b(min(lane & (C - 1), nm1));
    unsigned long long pE = ld_relaxed(Ein + min(lane & (C - 1), nm1));
    // Publishes the C bottom-edge strands caught in O during the chunk that
    // started at step t: O lanes [64-C, 64) hold columns t-64 .. t-64+C-1.
    // Every lane stores every time, so no path skips the store and the waits
    // for the prefetch loads can leave it in flight: lane L stores the word of
    // lane 64-C + L%C (fetched by one bpermute), so the C/64 copies of each
    // word carry the same value to the same address.  Columns outside [0, n)
    // (the first and last chunks of a band) go to trash words past the ring.
    // A chunk's words are published during the next chunk: publishing them
    // right after their chunk (the upper half of O copied down by
    // v_permlane32_swap instead of the bpermute) was measured slower, 146.1
    // against 144.4 ms at m=n=2^20.
    static_assert(C > 0 && (C & (C - 1)) == 0 && C <= kWave, "C must be a power of two <= 64");
    gword *trash = (gword *)(P.E + 2 * (size_t)n);
    // INSIDE: every column is known to be in [0, n), so the store needs no
    // range check (the interior chunks below).
    auto publish = [&](uint32_t Ox, int t, auto inside) {
        const int src = kWave - C + (lane & (C - 1));
        const uint32_t x = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)Ox);
        const int col = t + C - 2 * kWave + src;
        if constexpr (decltype(inside)::value) {
            st_relaxed(Eout + (uint32_t)col, tag_out | (unsigned long long)x);
        } else {
            const bool mine = col >= 0 && col < n;
            st_relaxed(mine ? Eout + col : trash + src, tag_out | (unsigned long long)x);
        }
    };
    // Turns the prefetched ring words of the chunk starting at t into lane 0's
    // conveyor I (waiting, bounded in time, until they carry this band's tag).
    // INSIDE: the chunk's columns are known to be < n.
    auto take_inputs = [&](int t, uint32_t &I, auto inside) -> bool {
        const int c = t + (lane & (C - 1));   // lanes >= C repeat lanes < C
        const bool need = lane < C && (decltype(inside)::value || c < n);
        unsigned long long e = pE;
        // keeps the tag test (and so the wait for the prefetched word) after
        // the chunk's last step: I still holds the conveyor the chunk left,
        // which that step wrote, so the compare cannot be hoisted next to
        // the load to wait there for the whole round trip
        asm volatile("" : "+v"(e) : "v"(I));
        bool ok = !need || (uint32_t)(e >> 32) == tag_in;
        if (__builtin_amdgcn_ballot_w64(!ok) != 0) {
            // The band above is not that far yet: poll.  Every round also
            // reads the abort flag and the clock; re-reading the ring words
            // back to back without them was measured slower.
            const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
            do {
                // a wave that shares its SIMD with other launches' waves
                // gives up its issue slots between polls
                if (A.poll_sleep) __builtin_amdgcn_s_sleep(2);
                if (__builtin_amdgcn_readfirstlane(ld_flag(A.abort_flag))) return false;
                if (__builtin_amdgcn_s_memrealtime() - t_start > A.timeout_ticks) {
                    raise_abort(A.abort_flag);
                    return false;
                }
                if (!ok) e = ld_relaxed(Ein + min(c, nm1));
                ok = !need || (uint32_t)(e >> 32) == tag_in;
            } while (__builtin_amdgcn_ballot_w64(!ok) != 0);
        }
        I = (uint32_t)e;
        return true;
    };
    uint32_t v = 0, Q = 0, Osave = 0;
    int32_t bq = 0, IB = pB;
    if (!take_inputs(0, Q, std::false_type{})) return false;
    // One chunk: prefetch chunk k+1, then publish chunk k-1 (the wait for the
    // loads at the end then leaves the store in flight), comb C steps, take
    // the next inputs.  The chunks run in three phases (ramp-in, interior,
    // ramp-out) so the hot loop holds only the unmasked steps and no
    // register copies between masked and unmasked paths; the interior loop
    // takes two chunks per iteration, which lets the compiler close it
    // without copying the h registers (with one chunk it copied them).
    // Issuing the prefetch halfway through the chunk instead was measured no
    // faster.
    using Masked = std::integral_constant<bool, true>;
    using Full = std::integral_constant<bool, false>;
    auto chunk = [&](auto colmask, int k) -> bool {
        const int t0 = k * C;
        if constexpr (S8) load_sym(psym, t0 + C);
        if constexpr (decltype(colmask)::value) {
            const int c = min(t0 + C + (lane & (C - 1)), nm1);
            if constexpr (!S8) pB = load_b(c);
            pE = ld_relaxed(Ein + c);
        } else {   // an interior chunk: the next chunk's columns are inside [0, n)
            const uint32_t c = (uint32_t)(t0 + C + (lane & (C - 1)));
            if constexpr (!S8) pB = load_b(c);
            pE = ld_relaxed(Ein + c);
        }
        publish(Osave, t0 - C, std::integral_constant<bool, !decltype(colmask)::value>{});
        if constexpr (S8)
            band_steps8<TM, C, ROWMASK, decltype(colmask)::value>(a, h, valid, v, Q, sym, t0, lane, n);
        else
            band_steps<TM, C, ROWMASK, decltype(colmask)::value, PK, OH>(a, h, valid, v, bq, Q, IB, t0, lane, n);
        Osave = Q;
        if (!decltype(colmask)::value || k + 1 < nchunks) {   // an interior chunk is never the last
            if constexpr (S8) {
#pragma unroll
                for (int q = 0; q < C / 4; ++q) sym[q] = psym[q];
            }
            IB = pB;
            if (!take_inputs(t0 + C, Q, std::integral_constant<bool, !decltype(colmask)::value>{})) return false;
        }
        return true;
    };
    // Interior chunks: no lane's column leaves [0, n) (t0 >= 63, t0 + C <= n),
    // the next chunk's prefetch stays inside (t0 + 2C <= n) and so does the
    // publish of the chunk before (columns t0 - 64 - C .. t0 - 65, so
    // t0 >= 64 + C).
    const int k_in = (kWave + C) / C;                      // first interior chunk
    const int k_out = max(k_in, min(n / C - 1, nchunks));  // first chunk after them
    int k = 0;
    for (; k < min(k_in, nchunks); ++k)
        if (!chunk(Masked{}, k)) return false;
    for (; k + 1 < k_out; k += 2)
        if (!chunk(Full{}, k) || !chunk(Full{}, k + 1)) return false;
    for (; k < k_out; ++k)
        if (!chunk(Full{}, k)) return false;
    for (; k < nchunks; ++k)
        if (!chunk(Masked{}, k)) return false;
    publish(Osave, (nchunks - 1) * C, std::false_type{});
    if constexpr (PK) {
#pragma unroll
        for (int k = 0; k < TM; ++k)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                if (r0 + k < P.m_[q]) P.h_out[q][P.m_[q] - 1 - (r0 + k)] = (int32_t)((h[k] >> (16 * q)) & 0xffffu);
    } else {
#pragma unroll
        for (int k = 0; k < TM; ++k)
            if (valid[k]) P.h_out[m - 1 - (r0 + k)] = (int32_t)h[k];
    }
    if (A.trace && t_id < A.trace_cap && lane == 0) A.trace[2 * (size_t)t_id + 1] = __builtin_amdgcn_s_memrealtime();
    return true;
}

template <int TM, int C>
__global__ __launch_bounds__(256) void comb_chain_kernel(ChainArgs A)
{
    const int lane = threadIdx.x & (kWave - 1);
    for (;;) {
        unsigned t = 0;
        if (lane == 0) t = atomicAdd(A.ticket, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if ((int)t >= A.ntasks) break;
        if (__builtin_amdgcn_readfirstlane(ld_flag(A.abort_flag))) break;
        const int2 task = A.tasks[t];
        const PairDesc P = A.pairs[task.x];
        const int g = task.y;
        const bool rowmask = (long long)(g + 1) * (kWave * TM) > (long long)P.m;
        // byte symbols when the pair has them and all of b fits in a byte
        const bool s8 = P.sym8 != nullptr && __builtin_amdgcn_readfirstlane(*P.sym8_bad) == 0;
        bool ok;
        if (s8)   // byte symbols pad a partial band's rows instead of masking them
            ok = band_pass<TM, C, false, false, PairDesc, true>(A, P, g, lane, (int)t);
        else
            ok = rowmask ? band_pass<TM, C, true, false>(A, P, g, lane, (int)t)
                         : band_pass<TM, C, false, false>(A, P, g, lane, (int)t);
        if (!ok) break;
    }
}

// The packed chain: ticket t runs band tasks[t].y of couple tasks[t].x.
template <int TM, int C>
__global__ __launch_bounds__(256) void comb_chain16_kernel(ChainArgs A)
{
    const int lane = threadIdx.x & (kWave - 1);
    for (;;) {
        unsigned t = 0;
        if (lane == 0) t = atomicAdd(A.ticket, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if ((int)t >= A.ntasks) break;
        if (__builtin_amdgcn_readfirstlane(ld_flag(A.abort_flag))) break;
        const int2 task = A.tasks[t];
        const CoupleDesc D = A.couples[task.x];
        const bool ok = __builtin_amdgcn_readfirstlane(D.onehot)
                            ? band_pass<TM, C, false, true, CoupleDesc, false, true>(A, D, task.y, lane, (int)t)
                            : band_pass<TM, C, false, true>(A, D, task.y, lane, (int)t);
        if (!ok) break;
    }
}

// ---- LDS group variant -----------------------------------------------------
//
// W consecutive bands of one pair share a workgroup, one wave per band (one
// per SIMD of a CU).  Inside the group the bottom-edge strands go from wave w
// to wave w+1 through an R-slot LDS ring guarded by two progress words (the
// producer's last published chunk, the consumer's last consumed chunk), and
// the b symbols are loaded once per group by wave 0 (coalesced, 64 B per
// chunk) into an LDS ring that every wave reads.  Only the group's first wave
// reads the global ring and only its last wave writes it.  Waves are not
// lock-stepped: each waits only for the data it needs and for a free slot,
// so the ring gives the same slack the global ring gives between bands,
// while a hand-off costs an LDS round trip instead of a global one.  A wave
// finishes chunk j-kShift's columns (kShift = 64/C: the lane skew) during its
// chunk j.  Every wait is bounded in time; an expired wait raises the global
// abort flag and an LDS stop word that every waiting wave checks.
constexpr int kXferSlots = 8;   // LDS ring between consecutive waves (chunks)
constexpr int kBRing = 64;      // b chunks staged in LDS

struct GroupLds {
    volatile int prog[4];   // prog[w]: last column chunk wave w published (-1: none)
    volatile int cons[4];   // cons[w]: last chunk wave w took from wave w-1
    volatile unsigned stop; // set by a wave whose wait expired
};
typedef __attribute__((address_space(3))) GroupLds lds_group;

// Spins until *word >= target (LDS), the stop word is set, or the time runs
// out (then raises both abort words).  Returns false when the wave must stop.
__device__ __forceinline__ bool lds_wait_ge(__attribute__((address_space(3))) volatile int *word, int target,
                                            lds_group *G, const ChainArgs &A)
{
    if (*word >= target) return true;
    const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
    while (*word < target) {
        if (G->stop) return false;
        if (__builtin_amdgcn_s_memrealtime() - t_start > A.timeout_ticks) {
            G->stop = 1u;
            raise_abort(A.abort_flag);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// ROLE bit 1: the group's first wave (reads the global ring, stages b);
// bit 2: its last wave (writes the global ring).  The role is a template
// argument so each wave's loop is straight-line code: with the role as a
// runtime branch the waits of the merged paths cover every outstanding load.
constexpr int kFirst = 1, kLast = 2;

template <int TM, int C, int W, bool ROWMASK, int ROLE>
__device__ bool group_band(const ChainArgs &A, const PairDesc &P, int g, int w, int lane,
                           uint32_t (*xfer_)[kXferSlots][C], int32_t (*bring_)[C], lds_group *G)
{
    constexpr int kShift = kWave / C;
    static_assert((W - 1) * (kShift + kXferSlots + 1) + 2 <= kBRing, "b ring too short for the group's spread");
    constexpr bool first = (ROLE & kFirst) != 0, last = (ROLE & kLast) != 0;
    // Every LDS access of the hand-off is volatile: the compiler keeps them
    // in program order and the LDS executes one wave's accesses in order, so
    // a progress word read before a data word vouches for that data.
    typedef __attribute__((address_space(3))) volatile uint32_t lds_u32;
    typedef __attribute__((address_space(3))) volatile int32_t lds_i32;
    lds_u32 *xfer_in = (lds_u32 *)(w > 0 ? &xfer_[w > 0 ? w - 1 : 0][0][0] : &xfer_[0][0][0]);
    lds_u32 *xfer_out = (lds_u32 *)&xfer_[last ? 0 : w][0][0];
    lds_i32 *bring = (lds_i32 *)&bring_[0][0];
    const int m = P.m, n = P.n;
    const int r0 = g * (kWave * TM) + lane * TM;
    int32_t a[TM];
    uint32_t h[TM];
    bool valid[TM];
    gsym *asym = (gsym *)P.a, *hin = (gsym *)P.h_in;
#pragma unroll
    for (int k = 0; k < TM; ++k) {
        const int r = r0 + k;
        valid[k] = ROWMASK ? (r < m) : true;
        a[k] = valid[k] ? asym[r] : 0;
        h[k] = valid[k] ? (hin ? (uint32_t)hin[m - 1 - r] : (uint32_t)(m - 1 - r)) : 0u;
    }
    // Let the row loads land here: left outstanding into the loop, the wait
    // for them would sit before the cells of every chunk and also cover the
    // last wave's in-flight ring store.
    __builtin_amdgcn_s_waitcnt(0);
    const uint32_t tag_in = (uint32_t)g;
    const unsigned long long tag_out = (unsigned long long)(g + 1) << 32;
    gcword *Ein = (gcword *)(P.E + (size_t)(g & 1) * (size_t)n);
    gword *Eout = (gword *)(P.E + (size_t)((g + 1) & 1) * (size_t)n);
    gword *trash = (gword *)(P.E + 2 * (size_t)n);
    gsym *bsym = (gsym *)P.b;
    const int nchunks = (n + kWave + C - 1) / C;
    const int nin = (n + C - 1) / C;   // chunks whose columns carry inputs
    const int nm1 = n - 1;
    const int ql = lane & (C - 1);
    // as in band_pass: every lane stores, C/64 copies of each word
    auto publish = [&](uint32_t Ox, int t) {
        const int src = kWave - C + ql;
        const uint32_t x = (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)Ox);
        const int col = t + C - 2 * kWave + src;
        st_relaxed(col >= 0 && col < n ? Eout + col : trash + src, tag_out | (unsigned long long)x);
    };
    unsigned long long pE = 0;
    auto take_inputs = [&](int t, uint32_t &I) -> bool {
        const int c = t + ql;
        const bool need = lane < C && c < n;
        unsigned long long e = pE;
        bool ok = !need || (uint32_t)(e >> 32) == tag_in;
        if (__builtin_amdgcn_ballot_w64(!ok) != 0) {
            // poll as band_pass does
            const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
            do {
                if (A.poll_sleep) __builtin_amdgcn_s_sleep(2);
                if (__builtin_amdgcn_readfirstlane(ld_flag(A.abort_flag)) || G->stop) return false;
                if (__builtin_amdgcn_s_memrealtime() - t_start > A.timeout_ticks) {
                    G->stop = 1u;
                    raise_abort(A.abort_flag);
                    return false;
                }
                if (!ok) e = ld_relaxed(Ein + min(c, nm1));
                ok = !need || (uint32_t)(e >> 32) == tag_in;
            } while (__builtin_amdgcn_ballot_w64(!ok) != 0);
        }
        I = (uint32_t)e;
        return true;
    };
    // chunk j's inputs from wave w-1 (wave >= 1): read now, checked later
    int32_t pn = 0;
    uint32_t In = 0;
    int32_t IBn = 0;
    auto lds_prefetch = [&](int j) {
        pn = G->prog[w - 1];
        In = xfer_in[(j & (kXferSlots - 1)) * C + ql];
        IBn = bring[(j & (kBRing - 1)) * C + ql];
    };
    auto lds_take = [&](int j, uint32_t &I, int32_t &IB) -> bool {
        if (__builtin_amdgcn_readfirstlane(pn) < j) {
            if (!lds_wait_ge(&G->prog[w - 1], j, G, A)) return false;
            In = xfer_in[(j & (kXferSlots - 1)) * C + ql];
            IBn = bring[(j & (kBRing - 1)) * C + ql];
        }
        I = In;
        IB = IBn;
        if (lane == 0) G->cons[w] = j;
        return true;
    };
    int32_t cn = 0;   // consumer word of wave w+1, read a chunk ahead

    uint32_t v = 0, Q = 0, Osave = 0;
    int32_t bq = 0, IB = 0, pB = 0;
    if constexpr (first) {
        pB = bsym[min(ql, nm1)];
        pE = ld_relaxed(Ein + min(ql, nm1));
        if (!take_inputs(0, Q)) return false;
        IB = pB;
        if (lane < C) bring[lane] = pB;
    } else {
        lds_prefetch(0);
        if (!lds_take(0, Q, IB)) return false;
    }
    for (int j = 0; j < nchunks; ++j) {
        const int t0 = j * C;
        // Order: global loads, then the ring store (so the wait for the
        // loads at the end of the chunk leaves the store in flight), then the
        // LDS prefetch (so the store's bpermute wait covers none of it).
        if constexpr (first) {
            if (j + 1 < nchunks) {
                const int c = min(t0 + C + ql, nm1);
                pB = bsym[c];
                pE = ld_relaxed(Ein + c);
            }
        }
        if constexpr (last) publish(Osave, t0 - C);
        if constexpr (!first) {
            if (j + 1 < nin) lds_prefetch(j + 1);
        }
        if constexpr (!last) cn = G->cons[w + 1];
        if (t0 >= kWave - 1 && t0 + C <= n)
            band_steps<TM, C, ROWMASK, false>(a, h, valid, v, bq, Q, IB, t0, lane, n);
        else
            band_steps<TM, C, ROWMASK, true>(a, h, valid, v, bq, Q, IB, t0, lane, n);
        Osave = Q;
        if constexpr (!last) {
            if (j >= kShift && j - kShift < nin) {
                const int jo = j - kShift;   // the column chunk O now completes
                if (__builtin_amdgcn_readfirstlane(cn) < jo - kXferSlots &&
                    !lds_wait_ge(&G->cons[w + 1], jo - kXferSlots, G, A))
                    return false;
                if (lane >= kWave - C) xfer_out[(jo & (kXferSlots - 1)) * C + lane - (kWave - C)] = Osave;
                if (lane == 0) G->prog[w] = jo;
            }
        }
        if constexpr (first) {
            if (j + 1 < nchunks) {
                IB = pB;
                if (!take_inputs(t0 + C, Q)) return false;
                if (lane < C) bring[((j + 1) & (kBRing - 1)) * C + lane] = pB;
            }
        } else {
            if (j + 1 < nin) {
                if (!lds_take(j + 1, Q, IB)) return false;
            } else {
                Q = 0;   // columns past n: no input
                IB = 0;
            }
        }
    }
    if constexpr (last) publish(Osave, (nchunks - 1) * C);
    gout *hout = (gout *)P.h_out;
#pragma unroll
    for (int k = 0; k < TM; ++k)
        if (valid[k]) hout[m - 1 - (r0 + k)] = (int32_t)h[k];
    return true;
}

// Ticket t runs band group tasks[t].y (bands W*y .. W*y+W-1) of pair
// tasks[t].x.  The only barriers are the three per task, which every wave
// passes whatever its band did.
template <int TM, int C, int W>
__global__ __launch_bounds__(kWave * W) void comb_group_kernel(ChainArgs A)
{
    static_assert(W >= 2 && W <= 4, "group of 2..4 waves");
    __shared__ uint32_t xfer[W - 1][kXferSlots][C];
    __shared__ int32_t bring[kBRing][C];
    __shared__ GroupLds group;
    __shared__ unsigned s_task, s_stop;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = (int)(threadIdx.x >> 6);
    lds_group *G = (lds_group *)&group;
    for (;;) {
        __syncthreads();  // every wave is done with the previous task's LDS
        if (threadIdx.x == 0) {
            const unsigned t = atomicAdd(A.ticket, 1u);
            s_task = t;
            s_stop = ((int)t >= A.ntasks || ld_flag(A.abort_flag)) ? 1u : 0u;
            for (int i = 0; i < 4; ++i) {
                G->prog[i] = -1;
                G->cons[i] = -1;
            }
            G->stop = 0u;
        }
        __syncthreads();
        const unsigned t = s_task, stop = s_stop;
        __syncthreads();
        if (stop) break;
        const int2 task = A.tasks[t];
        const PairDesc P = A.pairs[task.x];
        const int g0 = task.y * W;
        const int wact = min(W, P.nbands - g0);
        const int trace_id = (int)t * W + w;
        const bool traced = A.trace && trace_id < A.trace_cap && lane == 0;
        if (traced) A.trace[2 * (size_t)trace_id] = __builtin_amdgcn_s_memrealtime();
        if (w < wact) {
            const int g = g0 + w;
            // only a pair's last band has rows past m, and it is its group's last wave
            const bool rowmask = (long long)(g + 1) * (kWave * TM) > (long long)P.m;
            const int role = (w == 0 ? kFirst : 0) | (w == wact - 1 ? kLast : 0);
            bool ok;
            if (rowmask)
                ok = role == (kFirst | kLast)
                         ? group_band<TM, C, W, true, kFirst | kLast>(A, P, g, w, lane, xfer, bring, G)
                         : group_band<TM, C, W, true, kLast>(A, P, g, w, lane, xfer, bring, G);
            else if (role == (kFirst | kLast))
                ok = group_band<TM, C, W, false, kFirst | kLast>(A, P, g, w, lane, xfer, bring, G);
            else if (role == kFirst)
                ok = group_band<TM, C, W, false, kFirst>(A, P, g, w, lane, xfer, bring, G);
            else if (role == kLast)
                ok = group_band<TM, C, W, false, kLast>(A, P, g, w, lane, xfer, bring, G);
            else
                ok = group_band<TM, C, W, false, 0>(A, P, g, w, lane, xfer, bring, G);
            if (!ok && lane == 0) G->stop = 1u;  // let the group's other waves stop waiting
        }
        if (traced) A.trace[2 * (size_t)trace_id + 1] = __builtin_amdgcn_s_memrealtime();
    }
}

// Seeds the rings before a chain launch (blockIdx.y = pair): slot 0 holds
// the grid's top edge (v_in[c], or the ids m+c) under tag 0 for band 0,
// slot 1 holds a tag no band waits for.  A pair with no band (n == 0) gets
// its h_out here.  Block (0,0) clears the ticket and abort words, after
// folding the previous launch's abort flag into `sticky` (the context's
// status word, read and cleared by the host).  A kernel, not
// hipMemsetAsync, so the launch sequence can live inside a hipGraph.
__global__ void comb_prep_kernel(const PairDesc *pairs, unsigned *ctrl, unsigned *sticky)
{
    const PairDesc P = pairs[blockIdx.y];
    // 64-bit indices: a grid-stride step past a size near 2^31 must not wrap
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < P.n; c += stride) {
        const uint32_t v0 = P.v_in ? (uint32_t)P.v_in[c] : (uint32_t)(P.m + c);
        P.E[c] = (unsigned long long)v0;                       // tag 0
        P.E[(size_t)P.n + c] = 0xffffffff00000000ull;          // no band has this tag
    }
    if (P.nbands == 0)
        for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < P.m; r += stride)
            P.h_out[r] = P.h_in ? P.h_in[r] : (int32_t)r;
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        if (sticky && ctrl[1]) *sticky = 1u;
        ctrl[0] = 0u;
        ctrl[1] = 0u;
    }
}

// Seeds the rings of the packed chain (blockIdx.y = couple): column c of
// slot 0 carries both pairs' top-edge strands m + c (0xffff past a pair's
// n) under tag 0, slot 1 a tag no band waits for.  Block (0,0) clears the
// ticket and abort words after folding the abort flag into `sticky`.
__global__ void comb_prep16_kernel(const CoupleDesc *couples, unsigned *ctrl, unsigned *sticky)
{
    const CoupleDesc D = couples[blockIdx.y];
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < D.n; c += stride) {
        uint32_t v0 = 0;
        for (int q = 0; q < 2; ++q) v0 |= (c < D.n_[q] ? (uint32_t)(D.m_[q] + c) : 0xffffu) << (16 * q);
        D.E[c] = (unsigned long long)v0;                       // tag 0
        D.E[(size_t)D.n + c] = 0xffffffff00000000ull;          // no band has this tag
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        if (sticky && ctrl[1]) *sticky = 1u;
        ctrl[0] = 0u;
        ctrl[1] = 0u;
    }
}

// Both pairs' v_out from the last band's slot (blockIdx.y = couple).
__global__ void comb_extract16_kernel(const CoupleDesc *couples)
{
    const CoupleDesc D = couples[blockIdx.y];
    const unsigned long long *slot = D.E + (size_t)(D.nbands & 1) * (size_t)D.n;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < D.n; c += stride) {
        const uint32_t w = (uint32_t)slot[c];
        for (int q = 0; q < 2; ++q)
            if (c < D.n_[q]) D.v_out[q][c] = (int32_t)((w >> (16 * q)) & 0xffffu);
    }
}

// Clears n words (a kernel, so the launch sequence can be captured).
__global__ void lcsk_zero_kernel(int32_t *x, int32_t n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

// Byte symbols of b (the layout lcsk_sym8_stride describes): one thread per
// word of the four copies; copy 0 (which covers every column) raises *bad
// for a symbol outside [0, 255].
__global__ void comb_sym8_kernel(const int32_t *b, int32_t n, uint32_t *out, int32_t stride, int32_t *bad)
{
    const long long total = 4LL * stride;
    const long long step = (long long)gridDim.x * blockDim.x;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += step) {
        const int j = (int)(q / stride);
        const long long x = 4 * (q - (long long)j * stride) + j - kSym8Pad;   // column of byte 0
        uint32_t word = 0;
        bool out_of_range = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long c = x + i;
            if (c >= 0 && c < n) {
                const int32_t sym = b[c];
                out_of_range |= (uint32_t)sym > 255u;
                word |= ((uint32_t)sym & 0xffu) << (8 * i);
            }
        }
        out[q] = word;
        if (out_of_range && j == 0) atomicOr(bad, 1);
    }
}

// Range check of the packed chain's symbols (blockIdx.y = string): bit 0 of
// bad[item] for a symbol outside [0, 0xfffd], bit 1 for one above 15 (the
// one-hot cells need every symbol in [0, 15]).
__global__ void check16_kernel(const SymbolRange *ranges, int32_t *bad)
{
    const SymbolRange R = ranges[blockIdx.y];
    bool out = false, wide = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < R.len;
         i += (long long)gridDim.x * blockDim.x) {
        const uint32_t x = (uint32_t)R.x[i];
        out = out || x > 0xfffdu;
        wide = wide || x > 15u;
    }
    const int flags = (__builtin_amdgcn_ballot_w64(out) != 0 ? 1 : 0) | (__builtin_amdgcn_ballot_w64(wide) != 0 ? 2 : 0);
    if (flags && (threadIdx.x & 63) == 0) atomicOr(&bad[R.item], flags);
}

// v_out = the strands the last band left in its slot (blockIdx.y = pair).
__global__ void comb_extract_kernel(const PairDesc *pairs)
{
    const PairDesc P = pairs[blockIdx.y];
    const unsigned long long *slot = P.E + (size_t)(P.nbands & 1) * (size_t)P.n;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < P.n; c += stride)
        P.v_out[c] = (int32_t)(uint32_t)slot[c];
}

// row_to_col of each pair's permutation (blockIdx.y = pair): the strand that
// leaves row l on the right goes to column n+l, the one leaving column j at
// the bottom to column j (permutation_matrix.hpp:88-115).
__global__ void from_strands_kernel(const PairDesc *pairs, int32_t *perm, const int64_t *perm_off)
{
    const PairDesc P = pairs[blockIdx.y];
    int32_t *r2c = perm + perm_off[blockIdx.y];
    const long long stride = (long long)gridDim.x * blockDim.x;
    const uint32_t size = (uint32_t)(P.m + P.n);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)size; i += stride) {
        // ids are always < m+n after a completed comb; the test keeps an
        // aborted one (whose strands are undefined) from writing out of range
        const uint32_t id = (uint32_t)(i < P.m ? P.h_out[i] : P.v_out[i - P.m]);
        if (id < size) r2c[id] = (int32_t)(i < P.m ? P.n + i : i - P.m);
    }
}

// row_to_col of one pair's permutation from its device strands (FromStrands,
// src/permutation_matrix.hpp:88-115): h[l] -> n+l, v[j] -> j.
__global__ void r2c_from_strands_kernel(const int32_t *h, int32_t m, const int32_t *v, int32_t n, int32_t *r2c)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t size = (uint32_t)(m + n);
    if (i >= (long long)size) return;
    // ids are always < m+n after a completed comb; the test keeps an aborted
    // one (whose strands are undefined) from writing out of range
    const uint32_t id = (uint32_t)(i < m ? h[i] : v[i - m]);
    if (id < size) r2c[id] = (int32_t)(i < m ? n + i : i - m);
}

__global__ void invert_perm_kernel(const int32_t *r2c, int32_t size, int32_t *c2r)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < size && (uint32_t)r2c[i] < (uint32_t)size) c2r[r2c[i]] = (int32_t)i;
}

__global__ void strands_from_c2r_kernel(const int32_t *c2r, int32_t m, int32_t n, int32_t *h_out, int32_t *v_out)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v_out[i] = c2r[i];
    else if (i < (long long)n + m) h_out[i - n] = c2r[i];
}

__global__ void copy_kernel(int32_t *dst, const int32_t *src, long long count)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

// Identity strands for empty grids (m == 0 or n == 0).
__global__ void comb_identity_kernel(int32_t *h_out, int32_t m, int32_t *v_out, int32_t n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) h_out[i] = (int32_t)i;
    if (i < n) v_out[i] = (int32_t)(m + i);
}

}  // namespace lcsk

using namespace lcsk;

// Steps per chunk.  32 for both forms: against 16, half the per-chunk
// publish, prefetch and tag-check work outweighs the longer band-to-band lag
// (m=n=2^20 at TM=16: the register chain 143.9 against 145.6 ms, the LDS
// groups 154.7 against 166.0 ms; 8 and 64 steps are slower in both).
constexpr int kChainChunk = 32;
constexpr int kGroupChunk = 32;

template <int TM>
static int launch_chain_tm(const ChainArgs &args, int grid_blocks, int block_threads, hipStream_t stream)
{
    if (args.group_w == 1) {
        comb_chain_kernel<TM, kChainChunk><<<dim3(grid_blocks), dim3(block_threads), 0, stream>>>(args);
        return 0;
    }
    if (args.group_w == 4 && block_threads == 4 * kWave) {
        comb_group_kernel<TM, kGroupChunk, 4><<<dim3(grid_blocks), dim3(block_threads), 0, stream>>>(args);
        return 0;
    }
    return -22;
}

int lcsk_launch_chain(const ChainArgs &args, const PairDesc *d_pairs, int npairs, int max_n, int tile_m,
                      int grid_blocks, int block_threads, hipStream_t stream)
{
    if (npairs <= 0) return 0;
    const int fill_blocks = (int)std::min<long long>(1024, std::max<long long>(1, ((long long)max_n + 255) / 256));
    hipLaunchKernelGGL(comb_prep_kernel, dim3(fill_blocks, npairs), dim3(256), 0, stream, d_pairs, args.ticket,
                       args.sticky);
    if (args.ntasks > 0) {
        int rc;
        switch (tile_m) {
        case 1: rc = launch_chain_tm<1>(args, grid_blocks, block_threads, stream); break;
        case 2: rc = launch_chain_tm<2>(args, grid_blocks, block_threads, stream); break;
        case 4: rc = launch_chain_tm<4>(args, grid_blocks, block_threads, stream); break;
        case 8: rc = launch_chain_tm<8>(args, grid_blocks, block_threads, stream); break;
        case 16: rc = launch_chain_tm<16>(args, grid_blocks, block_threads, stream); break;
        case 32: rc = launch_chain_tm<32>(args, grid_blocks, block_threads, stream); break;
        default: rc = -22;
        }
        if (rc) return rc;
    }
    hipLaunchKernelGGL(comb_extract_kernel, dim3(fill_blocks, npairs), dim3(256), 0, stream, d_pairs);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_sym8(const int32_t *b, int32_t n, uint32_t *out, int32_t stride, int32_t *bad, hipStream_t stream)
{
    if (n <= 0 || stride < lcsk_sym8_stride(n)) return -22;
    hipLaunchKernelGGL(lcsk_zero_kernel, dim3(1), dim3(64), 0, stream, bad, 1);
    const long long words = 4LL * stride;
    const unsigned blocks = (unsigned)std::min<long long>((words + 255) / 256, 8192);
    hipLaunchKernelGGL(comb_sym8_kernel, dim3(blocks), dim3(256), 0, stream, b, n, out, stride, bad);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_check16(const SymbolRange *d_ranges, int32_t nranges, int32_t *d_bad, int32_t nitems,
                        hipStream_t stream)
{
    if (nranges <= 0) return 0;
    hipLaunchKernelGGL(lcsk_zero_kernel, dim3((unsigned)((nitems + 255) / 256)), dim3(256), 0, stream, d_bad, nitems);
    for (int32_t base = 0; base < nranges; base += 65535) {   // grid y is at most 65535
        const int32_t cnt = std::min<int32_t>(65535, nranges - base);
        hipLaunchKernelGGL(check16_kernel, dim3(4, (unsigned)cnt), dim3(256), 0, stream, d_ranges + base, d_bad);
    }
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_chain16(const ChainArgs &args, const CoupleDesc *d_couples, int ncouples, int max_n, int tile_m,
                        int grid_blocks, int block_threads, hipStream_t stream)
{
    if (ncouples <= 0) return 0;
    const int fill_blocks = (int)std::min<long long>(1024, std::max<long long>(1, ((long long)max_n + 255) / 256));
    hipLaunchKernelGGL(comb_prep16_kernel, dim3(fill_blocks, ncouples), dim3(256), 0, stream, d_couples, args.ticket,
                       args.sticky);
    if (args.ntasks > 0) {
        const dim3 grid(grid_blocks), block(block_threads);
        switch (tile_m) {
        case 1: comb_chain16_kernel<1, kChainChunk><<<grid, block, 0, stream>>>(args); break;
        case 2: comb_chain16_kernel<2, kChainChunk><<<grid, block, 0, stream>>>(args); break;
        case 4: comb_chain16_kernel<4, kChainChunk><<<grid, block, 0, stream>>>(args); break;
        case 8: comb_chain16_kernel<8, kChainChunk><<<grid, block, 0, stream>>>(args); break;
        case 16: comb_chain16_kernel<16, kChainChunk><<<grid, block, 0, stream>>>(args); break;
        case 32: comb_chain16_kernel<32, kChainChunk><<<grid, block, 0, stream>>>(args); break;
        default: return -22;
        }
    }
    hipLaunchKernelGGL(comb_extract16_kernel, dim3(fill_blocks, ncouples), dim3(256), 0, stream, d_couples);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_from_strands(const PairDesc *d_pairs, int npairs, int max_len, int32_t *perm,
                             const int64_t *d_perm_off, hipStream_t stream)
{
    if (npairs <= 0) return 0;
    const int blocks = (int)std::min<long long>(1024, std::max<long long>(1, ((long long)max_len + 255) / 256));
    hipLaunchKernelGGL(from_strands_kernel, dim3(blocks, npairs), dim3(256), 0, stream, d_pairs, perm, d_perm_off);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_r2c_from_strands(const int32_t *h, int32_t m, const int32_t *v, int32_t n, int32_t *r2c,
                                 hipStream_t stream)
{
    const long long size = (long long)m + n;
    if (size == 0) return 0;
    hipLaunchKernelGGL(r2c_from_strands_kernel, dim3((unsigned)((size + 255) / 256)), dim3(256), 0, stream, h, m, v, n,
                       r2c);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_strands_from_perm(const int32_t *r2c, int32_t m, int32_t n, int32_t *scratch, int32_t *h_out,
                                  int32_t *v_out, hipStream_t stream)
{
    const long long size = (long long)m + n;
    if (size == 0) return 0;
    const dim3 grid((unsigned)((size + 255) / 256));
    hipLaunchKernelGGL(invert_perm_kernel, grid, dim3(256), 0, stream, r2c, (int32_t)size, scratch);
    hipLaunchKernelGGL(strands_from_c2r_kernel, grid, dim3(256), 0, stream, scratch, m, n, h_out, v_out);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_copy(int32_t *dst, const int32_t *src, long long count, hipStream_t stream)
{
    if (count <= 0) return 0;
    const int blocks = (int)std::min<long long>(1024, (count + 255) / 256);
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, stream, dst, src, count);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}

int lcsk_launch_identity(int32_t *h_out, int32_t m, int32_t *v_out, int32_t n, hipStream_t stream)
{
    const long long mx = m > n ? m : n;
    if (mx == 0) return 0;
    hipLaunchKernelGGL(comb_identity_kernel, dim3((unsigned)((mx + 255) / 256)), dim3(256), 0, stream, h_out, m, v_out,
                       n);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}
