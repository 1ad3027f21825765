/*
 * CPU oracle for the semi-local LCS seaweed hot path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, as the checker or as the reported CPU baseline.  The
 * product (platform-independent-lcs_amd/) never links or calls it.
 *
 * Reference: egor-bv/platform-independent-lcs, paths relative to its root.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_comb_antidiag / orc_comb_rows / orc_comb_mt
 *       - LCS read-out equals the textbook DP (orc_llcs_dp) on random inputs;
 *       - every string-substring score H(i,j) equals the DP of a vs b[i:j);
 *       - hand-derived known answers (a={1,2}, b={1}; tiny all-match and
 *         no-match grids: identity permutation when nothing matches);
 *       - bijection on [0, m+n); a == b gives LCS = m;
 *       - the three orders agree with each other (order independence).
 *   orc_comb_simd  - equals orc_comb_rows bit for bit at every stripe edge
 *                    (8/16 lanes x 8 rows), column-block edge (2048), empty
 *                    grids, negative and 32-bit symbols, all-match/no-match,
 *                    for each instruction set the host has and 1..8 threads.
 *   orc_fixup      - combing the four quadrants in dependency order through
 *                    the fixup equals the whole comb.
 *   orc_from_strands / orc_hash
 *       - checked against the reference's own PermutationMatrix compiled
 *         from its source (oracle/_ref, built by oracle/Makefile) when that
 *         build is present, and against a Python big-integer restatement of
 *         the wrap-around arithmetic.
 *   orc_staggered_multiply / orc_steady_ant
 *       - composition law against the comb: P(a, b1 b2) equals the product
 *         glued on k = m, and P(a1 a2, b) equals the row-glued product on
 *         k = n, for random splits, alphabets and sizes;
 *       - orc_sticky_bruteforce (definition of the distance product) on
 *         random k <= 9 permutations.
 *   orc_llcs_dp    - textbook recurrence, checked on hand-computed cases.
 *   orc_llcs_bitpar - equals orc_llcs_dp on random pairs (alphabets 1..30,
 *                    negative and 32-bit symbols, word-boundary lengths
 *                    63/64/65/127/128/129, empty strings).
 *   orc_semilocal  - every query of all four quadrants equals the DP of the
 *                    corresponding substrings (small random pairs, empty
 *                    strings included); dominance counts equal numpy's count
 *                    of the definition; out-of-range queries give -1.
 */
#define _GNU_SOURCE
#include "lcs_oracle.h"

#include <sched.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ comb */

/* The cell rule, lcs_reference.hpp:47-56 / lcs_common.hpp:7-24. */
static inline void cell(int32_t sa, int32_t sb, int32_t *h, int32_t *v)
{
    int32_t hh = *h, vv = *v;
    int swap = (sa == sb) || (hh > vv);
    *h = swap ? vv : hh;
    *v = swap ? hh : vv;
}

/* lcs_reference.hpp:5-69 (Lcs_Semi_Reference): reverse a, init strands,
 * sweep the m+n-1 anti-diagonals. */
void orc_comb_antidiag(const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                       int32_t *h_out, int32_t *v_out)
{
    int32_t *ar = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    for (int32_t i = 0; i < m; ++i) ar[i] = a[m - 1 - i];
    for (int32_t i = 0; i < m; ++i) h_out[i] = i;
    for (int32_t j = 0; j < n; ++j) v_out[j] = m + j;
    int64_t diag_count = (int64_t)m + n - 1;
    for (int64_t d = 0; d < diag_count; ++d) {
        int32_t i_first = d < m ? (int32_t)(m - d - 1) : 0;
        int32_t j_first = d < m ? 0 : (int32_t)(d - m + 1);
        int32_t len = m - i_first < n - j_first ? m - i_first : n - j_first;
        for (int32_t s = 0; s < len; ++s)
            cell(ar[i_first + s], b[j_first + s], &h_out[i_first + s], &v_out[j_first + s]);
    }
    free(ar);
}

void orc_comb_rows(const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                   int32_t *h_out, int32_t *v_out)
{
    for (int32_t j = 0; j < n; ++j) v_out[j] = m + j;
    /* row i holds a[m-1-i]; v flows from row i+1 to row i */
    for (int32_t i = m - 1; i >= 0; --i) {
        int32_t sa = a[m - 1 - i];
        int32_t hh = i;
        for (int32_t j = 0; j < n; ++j) {
            int32_t vv = v_out[j];
            int swap = (sa == b[j]) | (hh > vv);
            int32_t nh = swap ? vv : hh;
            v_out[j] = swap ? hh : vv;
            hh = nh;
        }
        h_out[i] = hh;
    }
}

/* One block of the wavefront: rows [i_lo, i_hi) (processed from i_hi-1 down),
 * columns [j_lo, j_hi). */
static void comb_block(const int32_t *ar, const int32_t *b, int32_t *h, int32_t *v,
                       int32_t i_lo, int32_t i_hi, int32_t j_lo, int32_t j_hi)
{
    for (int32_t i = i_hi - 1; i >= i_lo; --i) {
        int32_t sa = ar[i];
        int32_t hh = h[i];
        for (int32_t j = j_lo; j < j_hi; ++j) {
            int32_t vv = v[j];
            int swap = (sa == b[j]) | (hh > vv);
            int32_t nh = swap ? vv : hh;
            v[j] = swap ? hh : vv;
            hh = nh;
        }
        h[i] = hh;
    }
}

void orc_comb_mt(const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                 int32_t *h_out, int32_t *v_out, int32_t threads)
{
    int32_t *ar = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    for (int32_t i = 0; i < m; ++i) ar[i] = a[m - 1 - i];
    for (int32_t i = 0; i < m; ++i) h_out[i] = i;
    for (int32_t j = 0; j < n; ++j) v_out[j] = m + j;
    if (threads < 1) threads = 1;
    /* blocks of RB rows x CB columns; narrow blocks keep >= 4 per thread on
     * an anti-diagonal when the grid is tall and thin */
    int32_t CB = n / (threads * 4);
    if (CB > 4096) CB = 4096;
    if (CB < 128) CB = 128;
    int32_t RB = m / (threads * 4);
    if (RB > 512) RB = 512;
    if (RB < 64) RB = 64;
    int32_t nrb = (m + RB - 1) / RB, ncb = (n + CB - 1) / CB;
    /* block row I covers the I-th group of rows counted from the top of the
     * grid (high i); its v input comes from block row I-1. */
    for (int32_t d = 0; d + 1 <= nrb + ncb - 1 && nrb > 0 && ncb > 0; ++d) {
        int32_t I0 = d < nrb ? d : nrb - 1;
        int32_t J0 = d - I0;
        int32_t cnt = (I0 + 1 < ncb - J0) ? I0 + 1 : ncb - J0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (int32_t s = 0; s < cnt; ++s) {
            int32_t I = I0 - s, J = J0 + s;
            int32_t i_hi = m - I * RB, i_lo = i_hi - RB < 0 ? 0 : i_hi - RB;
            int32_t j_lo = J * CB, j_hi = j_lo + CB > n ? n : j_lo + CB;
            comb_block(ar, b, h_out, v_out, i_lo, i_hi, j_lo, j_hi);
        }
    }
    free(ar);
}

/* ------------------------------------------------------ vectorised comb */
/*
 * orc_comb_simd: the reference's tiled multithreaded comb on CPU SIMD lanes
 * (Lcs_Semi_Tiled_MT, lcs_tiled.hpp:95-137; update_stripe,
 * lcs_common.hpp:131-173, whose sub-group lanes each own TILE_M rows).
 *
 * Grid orientation: row r holds a[r] (the reference's reversed row m-1-r),
 * v strands flow from row r-1 into row r, row r starts with strand m-1-r and
 * column c with m+c.  A stripe is R = W*TM consecutive rows; lane l keeps
 * rows l*TM .. l*TM+TM-1 of it (symbols and h strands) in registers.  At
 * step t lane l combs column t-l, top row first; the v strand then moves one
 * lane down (lane shift), lane 0 takes the stripe's input v of column t and
 * lane W-1 hands the stripe's output for column t-W+1 on.  v lives in a
 * column-reversed array (vrev[P + n-1-c]), b likewise, so lane l's symbol at
 * step t is one contiguous load.  Rows past m never match and hold strand 0,
 * the smallest id, so every v strand passes them unchanged.
 *
 * Stripes are dealt round-robin to the threads; a thread combs its stripe in
 * column blocks of CB, each after the stripe above has finished that block
 * (a per-stripe progress counter, release/acquire).  Ramp steps at a block's
 * two ends mask the lanes whose column lies outside the block.
 */
#include <immintrin.h>
#include <stdatomic.h>

#define SIMD_TM 8
#define SIMD_PAD 64
#define SIMD_CB 2048

typedef struct {
    const int32_t *at;      /* [stripe][k][lane] symbols                    */
    uint32_t *ht;           /* [stripe][k][lane] h strands                  */
    const int32_t *brev;    /* brev[SIMD_PAD + n-1-c] = b[c]                */
    uint32_t *vrev;         /* vrev[SIMD_PAD + n-1-c] = v of column c       */
    int32_t m, n, nst, W;
} simd_grid;

/* active lanes of a step: j0 <= t - l < j1 */
static inline uint32_t ramp_mask(int64_t t, int64_t j0, int64_t j1, int W)
{
    int64_t lo = t - j1 + 1, hi = t - j0;
    if (lo < 0) lo = 0;
    if (hi > W - 1) hi = W - 1;
    if (hi < lo) return 0;
    return (uint32_t)((((uint64_t)1 << (hi + 1)) - 1) & ~(((uint64_t)1 << lo) - 1));
}

__attribute__((target("avx512f"))) static void
stripe_block_avx512(const simd_grid *G, int32_t I, int32_t j0, int32_t j1)
{
    enum { W = 16, TM = SIMD_TM };
    const int32_t n = G->n;
    const int32_t *at = G->at + (size_t)I * TM * W;
    uint32_t *ht = G->ht + (size_t)I * TM * W;
    __m512i A[TM], H[TM];
    __mmask16 valid[TM];
    const int rowmask = (int64_t)(I + 1) * W * TM > G->m;
#pragma GCC unroll 16
    for (int k = 0; k < TM; ++k) {
        A[k] = _mm512_loadu_si512((const void *)(at + k * W));
        H[k] = _mm512_loadu_si512((const void *)(ht + k * W));
        uint32_t vm = 0;
        for (int l = 0; l < W; ++l) vm |= (uint32_t)((int64_t)I * W * TM + l * TM + k < G->m) << l;
        valid[k] = (__mmask16)vm;
    }
    const __m512i ones = _mm512_set1_epi32(-1);
    __m512i V = _mm512_setzero_si512();
    const int32_t *bp = G->brev + SIMD_PAD + n - 1;
    uint32_t *vp = G->vrev + SIMD_PAD + n - 1;
    const int64_t t_end = (int64_t)j1 + W - 1;
    for (int64_t t = j0; t < t_end; ++t) {
        const uint32_t vin = t < j1 ? vp[-t] : 0u;
        V = _mm512_alignr_epi32(V, _mm512_set1_epi32((int)vin), W - 1);
        const __m512i B = _mm512_loadu_si512((const void *)(bp - t));
        const int full = t >= (int64_t)j0 + W - 1 && t < j1;
        if (full && !rowmask) {
#pragma GCC unroll 16
            for (int k = 0; k < TM; ++k) {
                const __mmask16 mt = _mm512_cmpeq_epi32_mask(A[k], B);
                const __m512i hn = _mm512_min_epu32(_mm512_mask_mov_epi32(H[k], mt, ones), V);
                V = _mm512_ternarylogic_epi32(H[k], V, hn, 0x96);
                H[k] = hn;
            }
        } else {
            const __mmask16 act = (__mmask16)ramp_mask(t, j0, j1, W);
#pragma GCC unroll 16
            for (int k = 0; k < TM; ++k) {
                const __mmask16 mt = _mm512_cmpeq_epi32_mask(A[k], B) & valid[k];
                const __m512i hn = _mm512_min_epu32(_mm512_mask_mov_epi32(H[k], mt, ones), V);
                V = _mm512_ternarylogic_epi32(H[k], V, hn, 0x96);
                H[k] = _mm512_mask_mov_epi32(H[k], act, hn);
            }
        }
        /* lane W-1 leaves the stripe with column t-W+1 */
        if (t - (W - 1) >= j0 && t - (W - 1) < j1) _mm512_mask_storeu_epi32((void *)(vp - t), (__mmask16)0x8000, V);
    }
#pragma GCC unroll 16
    for (int k = 0; k < TM; ++k) _mm512_storeu_si512((void *)(ht + k * W), H[k]);
}

__attribute__((target("avx2"))) static void
stripe_block_avx2(const simd_grid *G, int32_t I, int32_t j0, int32_t j1)
{
    enum { W = 8, TM = SIMD_TM };
    const int32_t n = G->n;
    const int32_t *at = G->at + (size_t)I * TM * W;
    uint32_t *ht = G->ht + (size_t)I * TM * W;
    __m256i A[TM], H[TM], valid[TM];
#pragma GCC unroll 16
    for (int k = 0; k < TM; ++k) {
        A[k] = _mm256_loadu_si256((const __m256i *)(at + k * W));
        H[k] = _mm256_loadu_si256((const __m256i *)(ht + k * W));
        int32_t vm[8];
        for (int l = 0; l < W; ++l) vm[l] = (int64_t)I * W * TM + l * TM + k < G->m ? -1 : 0;
        valid[k] = _mm256_loadu_si256((const __m256i *)vm);
    }
    const int rowmask = (int64_t)(I + 1) * W * TM > G->m;
    const __m256i rot = _mm256_setr_epi32(7, 0, 1, 2, 3, 4, 5, 6);
    __m256i V = _mm256_setzero_si256();
    const int32_t *bp = G->brev + SIMD_PAD + n - 1;
    uint32_t *vp = G->vrev + SIMD_PAD + n - 1;
    const int64_t t_end = (int64_t)j1 + W - 1;
    for (int64_t t = j0; t < t_end; ++t) {
        const uint32_t vin = t < j1 ? vp[-t] : 0u;
        V = _mm256_blend_epi32(_mm256_permutevar8x32_epi32(V, rot), _mm256_set1_epi32((int)vin), 0x01);
        const __m256i B = _mm256_loadu_si256((const __m256i *)(bp - t));
        const int full = t >= (int64_t)j0 + W - 1 && t < j1;
        if (full && !rowmask) {
#pragma GCC unroll 16
            for (int k = 0; k < TM; ++k) {
                const __m256i hx = _mm256_or_si256(H[k], _mm256_cmpeq_epi32(A[k], B));
                const __m256i hn = _mm256_min_epu32(hx, V);
                V = _mm256_xor_si256(_mm256_xor_si256(H[k], V), hn);
                H[k] = hn;
            }
        } else {
            const uint32_t am = ramp_mask(t, j0, j1, W);
            int32_t act_l[8];
            for (int l = 0; l < W; ++l) act_l[l] = (am >> l) & 1 ? -1 : 0;
            const __m256i act = _mm256_loadu_si256((const __m256i *)act_l);
#pragma GCC unroll 16
            for (int k = 0; k < TM; ++k) {
                const __m256i mt = _mm256_and_si256(_mm256_cmpeq_epi32(A[k], B), valid[k]);
                const __m256i hn = _mm256_min_epu32(_mm256_or_si256(H[k], mt), V);
                V = _mm256_xor_si256(_mm256_xor_si256(H[k], V), hn);
                H[k] = _mm256_blendv_epi8(H[k], hn, act);
            }
        }
        if (t - (W - 1) >= j0 && t - (W - 1) < j1) vp[-(t - (W - 1))] = (uint32_t)_mm256_extract_epi32(V, 7);
    }
#pragma GCC unroll 16
    for (int k = 0; k < TM; ++k) _mm256_storeu_si256((__m256i *)(ht + k * W), H[k]);
}

int32_t orc_comb_simd(const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                      int32_t *h_out, int32_t *v_out, int32_t threads, int32_t isa)
{
    __builtin_cpu_init();
    const int has512 = __builtin_cpu_supports("avx512f"), has2 = __builtin_cpu_supports("avx2");
    if (isa == 0) isa = has512 ? 2 : has2 ? 1 : 3;
    if ((isa == 2 && !has512) || (isa == 1 && !has2) || isa < 1 || isa > 3) return -1;
    if (threads < 1) threads = 1;
    if (isa == 3 || m == 0 || n == 0) {
        orc_comb_mt(a, m, b, n, h_out, v_out, threads);
        return isa;
    }
    const int W = isa == 2 ? 16 : 8, TM = SIMD_TM, R = W * TM;
    const int32_t nst = (int32_t)(((int64_t)m + R - 1) / R);
    const size_t cells = (size_t)nst * R;
    int32_t *at = (int32_t *)malloc(sizeof(int32_t) * cells);
    uint32_t *ht = (uint32_t *)malloc(sizeof(uint32_t) * cells);
    int32_t *brev = (int32_t *)calloc((size_t)n + 2 * SIMD_PAD, sizeof(int32_t));
    uint32_t *vrev = (uint32_t *)calloc((size_t)n + 2 * SIMD_PAD, sizeof(uint32_t));
    const int32_t ncb = (int32_t)(((int64_t)n + SIMD_CB - 1) / SIMD_CB);
    atomic_int *progress = (atomic_int *)malloc(sizeof(atomic_int) * (size_t)nst);
    for (int32_t I = 0; I < nst; ++I)
        for (int k = 0; k < TM; ++k)
            for (int l = 0; l < W; ++l) {
                const int64_t r = (int64_t)I * R + (int64_t)l * TM + k;
                const size_t x = ((size_t)I * TM + k) * W + l;
                at[x] = r < m ? a[r] : 0;
                ht[x] = r < m ? (uint32_t)(m - 1 - r) : 0u;
            }
    for (int32_t c = 0; c < n; ++c) {
        brev[SIMD_PAD + n - 1 - c] = b[c];
        vrev[SIMD_PAD + n - 1 - c] = (uint32_t)(m + c);
    }
    for (int32_t I = 0; I < nst; ++I) atomic_init(&progress[I], 0);
    const simd_grid G = {at, ht, brev, vrev, m, n, nst, W};
#pragma omp parallel num_threads(threads)
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        for (int32_t I = tid; I < nst; I += nth)
            for (int32_t J = 0; J < ncb; ++J) {
                if (I > 0) {
                    long spins = 0;
                    while (atomic_load_explicit(&progress[I - 1], memory_order_acquire) <= J) {
                        _mm_pause();
                        if (++spins > 4096) { sched_yield(); spins = 0; }
                    }
                }
                const int32_t j0 = J * SIMD_CB, j1 = j0 + SIMD_CB < n ? j0 + SIMD_CB : n;
                if (isa == 2) stripe_block_avx512(&G, I, j0, j1);
                else stripe_block_avx2(&G, I, j0, j1);
                atomic_store_explicit(&progress[I], J + 1, memory_order_release);
            }
    }
    for (int32_t I = 0; I < nst; ++I)
        for (int k = 0; k < TM; ++k)
            for (int l = 0; l < W; ++l) {
                const int64_t r = (int64_t)I * R + (int64_t)l * TM + k;
                if (r < m) h_out[m - 1 - r] = (int32_t)ht[((size_t)I * TM + k) * W + l];
            }
    for (int32_t c = 0; c < n; ++c) v_out[c] = (int32_t)vrev[SIMD_PAD + n - 1 - c];
    free(at);
    free(ht);
    free(brev);
    free(vrev);
    free(progress);
    return isa;
}

/* lcs_residual_fixup.hpp:74-122 (Lcs_Semi_Fixup). */
void orc_fixup(const int32_t *a_prepared, const int32_t *b_prepared,
               int32_t *h_strands, int32_t *v_strands,
               int32_t i0, int32_t isize, int32_t j0, int32_t jsize)
{
    if (isize == 0 || jsize == 0) return;
    int32_t *h = h_strands + i0, *v = v_strands + j0;
    const int32_t *ap = a_prepared + i0, *bp = b_prepared + j0;
    int64_t diag_count = (int64_t)isize + jsize - 1;
    for (int64_t d = 0; d < diag_count; ++d) {
        int32_t i_first = d < isize ? (int32_t)(isize - d - 1) : 0;
        int32_t j_first = d < isize ? 0 : (int32_t)(d - isize + 1);
        int32_t len = isize - i_first < jsize - j_first ? isize - i_first : jsize - j_first;
        for (int32_t s = 0; s < len; ++s)
            cell(ap[i_first + s], bp[j_first + s], &h[i_first + s], &v[j_first + s]);
    }
}

/* ------------------------------------------------------------ read-out */

/* permutation_matrix.hpp:88-115 */
void orc_from_strands(const int32_t *h, int32_t m, const int32_t *v, int32_t n,
                      int32_t *row_to_col, int32_t *col_to_row)
{
    int32_t size = m + n;
    for (int32_t i = 0; i < size; ++i) row_to_col[i] = col_to_row[i] = -1;
    for (int32_t l = 0; l < m; ++l)
        if (0 <= h[l] && h[l] < size) { row_to_col[h[l]] = n + l; col_to_row[n + l] = h[l]; }
    for (int32_t r = m; r < size; ++r) {
        int32_t s = v[r - m];
        if (0 <= s && s < size) { row_to_col[s] = r - m; col_to_row[r - m] = s; }
    }
}

/* permutation_matrix.hpp:192-203, R = 4294967279, M = 4294967291 */
int64_t orc_hash(const int32_t *col_to_row, int32_t size)
{
    const uint64_t R = 4294967279ull;
    const int64_t M = 4294967291ll;
    int64_t res = 0;
    for (int32_t c = 0; c < size; ++c) {
        uint64_t t = R * (uint64_t)res + (uint64_t)(int64_t)col_to_row[c];
        res = (int64_t)t % M;
    }
    return res;
}

int32_t orc_llcs_from_strands(const int32_t *v_out, int32_t m, int32_t n)
{
    int32_t c = 0;
    for (int32_t j = 0; j < n; ++j) c += v_out[j] < m;
    return c;
}

void orc_string_substring(const int32_t *row_to_col, int32_t m, int32_t n, int32_t *out)
{
    int32_t W = n + 1;
    int32_t *cnt = (int32_t *)calloc((size_t)W, sizeof(int32_t));
    memset(out, 0, sizeof(int32_t) * (size_t)W * W);
    /* cnt[j] = #{ top strands starting in column >= i that end at a bottom
     * column < j } */
    for (int32_t i = n; i >= 0; --i) {
        if (i < n) {
            int32_t col = row_to_col[m + i];
            if (col >= 0 && col < n)
                for (int32_t j = col + 1; j <= n; ++j) cnt[j]++;
        }
        for (int32_t j = i; j <= n; ++j) out[i * W + j] = (j - i) - cnt[j];
    }
    free(cnt);
}

static int cmp_i32(const void *x, const void *y)
{
    const int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

/* lcs_reference.hpp:135-211 (the disabled binary bit-parallel LLCS) packs
 * strands into words; this is the word-parallel recurrence for any alphabet. */
int32_t orc_llcs_bitpar(const int32_t *a, int32_t m, const int32_t *b, int32_t n)
{
    if (m <= 0 || n <= 0) return 0;
    const int64_t W = ((int64_t)m + 63) / 64;
    int32_t *sym = (int32_t *)malloc(sizeof(int32_t) * (size_t)m);
    memcpy(sym, a, sizeof(int32_t) * (size_t)m);
    qsort(sym, (size_t)m, sizeof(int32_t), cmp_i32);
    int64_t k = 0;
    for (int64_t i = 0; i < m; ++i)
        if (i == 0 || sym[i] != sym[i - 1]) sym[k++] = sym[i];
    uint64_t *M = (uint64_t *)calloc((size_t)(k * W), sizeof(uint64_t));
    for (int64_t i = 0; i < m; ++i) {
        const int32_t *p = (const int32_t *)bsearch(&a[i], sym, (size_t)k, sizeof(int32_t), cmp_i32);
        M[(p - sym) * W + i / 64] |= 1ull << (i % 64);
    }
    uint64_t *V = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)W);
    for (int64_t w = 0; w < W; ++w) V[w] = ~0ull;
    for (int32_t j = 0; j < n; ++j) {
        const int32_t *p = (const int32_t *)bsearch(&b[j], sym, (size_t)k, sizeof(int32_t), cmp_i32);
        if (!p) continue;   /* matches nothing: V unchanged */
        const uint64_t *Mc = M + (p - sym) * W;
        uint64_t carry = 0;
        for (int64_t w = 0; w < W; ++w) {
            const uint64_t v = V[w], u = v & Mc[w];
            const uint64_t s1 = v + u, s = s1 + carry;
            carry = (uint64_t)(s1 < v) | (uint64_t)(s < s1);
            V[w] = s | (v & ~Mc[w]);
        }
    }
    int64_t zeros = 0;
    for (int64_t w = 0; w < W; ++w) {
        uint64_t valid = (w == W - 1 && m % 64) ? ((1ull << (m % 64)) - 1) : ~0ull;
        zeros += __builtin_popcountll(~V[w] & valid);
    }
    free(sym);
    free(M);
    free(V);
    return (int32_t)zeros;
}

static int32_t dominance_count(const int32_t *p, int32_t size, int64_t r0, int64_t c0)
{
    int32_t cnt = 0;
    for (int64_t r = r0 < 0 ? 0 : r0; r < size; ++r) cnt += (int64_t)p[r] < c0;
    return cnt;
}

static int32_t seaweed_h(const int32_t *p, int32_t m, int32_t n, int64_t i, int64_t j)
{
    return (int32_t)(j - i) - dominance_count(p, m + n, m + i, j);
}

void orc_semilocal(const int32_t *row_to_col, int32_t m, int32_t n, int32_t kind,
                   const int32_t *q0, const int32_t *q1, int32_t nq, int32_t *out)
{
    for (int32_t q = 0; q < nq; ++q) {
        const int64_t x = q0[q], y = q1[q];
        int32_t r = -1;
        switch (kind) {
        case 0: if (0 <= x && x <= y && y <= n) r = seaweed_h(row_to_col, m, n, x, y); break;
        case 1: if (0 <= x && x <= y && y <= m) r = seaweed_h(row_to_col, m, n, -x, n + (m - y)) - (int32_t)x - (int32_t)(m - y); break;
        case 2: if (0 <= x && x <= m && 0 <= y && y <= n) r = seaweed_h(row_to_col, m, n, y, n + (m - x)) - (int32_t)(m - x); break;
        case 3: if (0 <= x && x <= m && 0 <= y && y <= n) r = seaweed_h(row_to_col, m, n, -x, y) - (int32_t)x; break;
        case 4: r = dominance_count(row_to_col, m + n, x, y); break;
        }
        out[q] = r;
    }
}

int32_t orc_llcs_dp(const int32_t *a, int32_t m, const int32_t *b, int32_t n)
{
    int32_t *prev = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    int32_t *cur = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
    for (int32_t i = 1; i <= m; ++i) {
        cur[0] = 0;
        for (int32_t j = 1; j <= n; ++j) {
            int32_t best = prev[j] > cur[j - 1] ? prev[j] : cur[j - 1];
            if (a[i - 1] == b[j - 1] && prev[j - 1] + 1 > best) best = prev[j - 1] + 1;
            cur[j] = best;
        }
        int32_t *t = prev; prev = cur; cur = t;
    }
    int32_t r = prev[n];
    free(prev); free(cur);
    return r;
}

/* -------------------------------------------------------- steady ant */

typedef struct { int32_t *r2c, *c2r; int32_t size; } perm_t;

static void perm_set(perm_t *p, int32_t r, int32_t c) { p->r2c[r] = c; p->c2r[c] = r; }
static void perm_unset(perm_t *p)
{
    for (int32_t i = 0; i < p->size; ++i) p->r2c[i] = p->c2r[i] = -1;
}
static perm_t perm_alloc(int32_t size)
{
    perm_t p;
    p.size = size;
    p.r2c = (int32_t *)malloc(sizeof(int32_t) * (size_t)(size > 0 ? size : 1));
    p.c2r = (int32_t *)malloc(sizeof(int32_t) * (size_t)(size > 0 ? size : 1));
    perm_unset(&p);
    return p;
}
static void perm_free(perm_t *p) { free(p->r2c); free(p->c2r); }

/* braid_multiplication.hpp:5-19: points of p with col in [c0, c1), rows
 * renumbered in order. */
static void vert_slice(const perm_t *p, int32_t c0, int32_t c1, int32_t *row_map, perm_t *out)
{
    int32_t idx = 0;
    for (int32_t r = 0; r < p->size; ++r) {
        int32_t c = p->r2c[r];
        if (c >= c0 && c < c1) { row_map[idx] = r; perm_set(out, idx, c - c0); ++idx; }
    }
}

/* braid_multiplication.hpp:22-36: points of q with row in [r0, r1), columns
 * renumbered in order. */
static void horz_slice(const perm_t *q, int32_t r0, int32_t r1, int32_t *col_map, perm_t *out)
{
    int32_t idx = 0;
    for (int32_t c = 0; c < q->size; ++c) {
        int32_t r = q->c2r[c];
        if (r >= r0 && r < r1) { col_map[idx] = c; perm_set(out, r - r0, idx); ++idx; }
    }
}

/* Incremental dominance sums, braid_multiplication.hpp:39-119.
 *   hi-sum(dr, dc) = #{ (r,c) in r_hi : r <  dr, c <  dc }
 *   lo-sum(dr, dc) = #{ (r,c) in r_lo : r >= dr, c >= dc }        */
static int32_t hi_left(int32_t row, int32_t col, int32_t s, const perm_t *p)
{ if (col == 0) return s; int32_t rc = p->c2r[col - 1]; if (rc != -1 && rc < row) s--; return s; }
static int32_t hi_right(int32_t row, int32_t col, int32_t s, const perm_t *p)
{ if (col >= p->size) return 0; int32_t rc = p->c2r[col]; if (rc != -1 && rc < row) s++; return s; }
static int32_t hi_up(int32_t row, int32_t col, int32_t s, const perm_t *p)
{ if (row == 0) return s; int32_t cc = p->r2c[row - 1]; if (cc != -1 && cc < col) s--; return s; }
static int32_t hi_down(int32_t row, int32_t col, int32_t s, const perm_t *p)
{ if (row >= p->size) return 0; int32_t cc = p->r2c[row]; if (cc != -1 && cc < col) s++; return s; }
static int32_t lo_left(int32_t row, int32_t col, int32_t s, const perm_t *p)
{ if (col == 0) return s; int32_t rc = p->c2r[col - 1]; if (rc >= row && rc != -1) s++; return s; }
static int32_t lo_right(int32_t row, int32_t col, int32_t s, const perm_t *p)
{ if (col >= p->size) return 0; int32_t rc = p->c2r[col]; if (rc >= row && rc != -1) s--; return s; }
static int32_t lo_up(int32_t row, int32_t col, int32_t s, const perm_t *p)
{ if (row == 0) return s; int32_t cc = p->r2c[row - 1]; if (cc >= col && cc != -1) s++; return s; }
static int32_t lo_down(int32_t row, int32_t col, int32_t s, const perm_t *p)
{ if (row >= p->size) return 0; int32_t cc = p->r2c[row]; if (cc >= col && cc != -1) s--; return s; }

/* braid_multiplication.hpp:123-189: walk from (n, -1) to the top/right edge
 * along the boundary of { hi-sum - lo-sum >= 0 }, collecting good points. */
static int32_t ant_passage(const perm_t *r_lo, const perm_t *r_hi, int32_t n,
                           int32_t *good_row, int32_t *good_col)
{
    int32_t ngood = 0;
    int32_t cur_row = r_hi->size, cur_col = -1;
    int32_t rhi = 0, rlo = 0;
    int went_right = 0;
    for (;;) {
        if (cur_row == 0) break;
        if (cur_col == n) break;
        int32_t dr = cur_row - 1, dc = cur_col + 1;
        if (went_right) {
            rhi = hi_right(dr, dc - 1, rhi, r_hi);
            rlo = lo_right(dr, dc - 1, rlo, r_lo);
        } else {
            rhi = hi_up(dr + 1, dc, rhi, r_hi);
            rlo = lo_up(dr + 1, dc, rlo, r_lo);
        }
        if (rhi - rlo < 0) { went_right = 1; cur_col++; }
        else               { went_right = 0; cur_row--; }   /* rhi - rlo == 0 */
        if (dc > 0) {
            int32_t d_tl = hi_left(dr, dc, rhi, r_hi) - lo_left(dr, dc, rlo, r_lo);
            int32_t d_br = hi_down(dr, dc, rhi, r_hi) - lo_down(dr, dc, rlo, r_lo);
            if (d_tl < 0 && d_br > 0) { good_row[ngood] = dr; good_col[ngood] = dc - 1; ++ngood; }
        }
    }
    return ngood;
}

/* braid_multiplication.hpp:191-202 */
static void inverse_mapping(const perm_t *shr, const int32_t *row_map, const int32_t *col_map, perm_t *flat)
{
    perm_unset(flat);
    for (int32_t c = 0; c < shr->size; ++c) {
        int32_t r = shr->c2r[c];
        perm_set(flat, row_map[r], col_map[c]);
    }
}

/* braid_multiplication.hpp:204-270 (steady_ant_seq); result left in p. */
static void steady_ant_rec(perm_t *p, perm_t *q)
{
    int32_t n = p->size;
    if (n <= 1) return;
    int32_t split = n / 2, lo = split, hi = n - split;
    perm_t p_lo = perm_alloc(lo), q_lo = perm_alloc(lo), p_hi = perm_alloc(hi), q_hi = perm_alloc(hi);
    int32_t *maps = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * n));
    int32_t *p_lo_rows = maps, *q_lo_cols = maps + lo, *p_hi_rows = maps + 2 * lo, *q_hi_cols = maps + 2 * lo + hi;
    vert_slice(p, 0, split, p_lo_rows, &p_lo);
    horz_slice(q, 0, split, q_lo_cols, &q_lo);
    vert_slice(p, split, n, p_hi_rows, &p_hi);
    horz_slice(q, split, q->size, q_hi_cols, &q_hi);
    steady_ant_rec(&p_lo, &q_lo);
    steady_ant_rec(&p_hi, &q_hi);
    perm_t *r_lo = p, *r_hi = q;
    inverse_mapping(&p_lo, p_lo_rows, q_lo_cols, r_lo);
    inverse_mapping(&p_hi, p_hi_rows, q_hi_cols, r_hi);
    int32_t *good = (int32_t *)malloc(sizeof(int32_t) * (size_t)(4 * n + 4));
    int32_t ng = ant_passage(r_lo, r_hi, n, good, good + 2 * n + 2);
    for (int32_t i = 0; i < n; ++i) {
        int32_t c = r_hi->r2c[i];
        if (c != -1) perm_set(r_lo, i, c);
    }
    for (int32_t g = 0; g < ng; ++g) perm_set(r_lo, good[g], good[2 * n + 2 + g]);
    free(good); free(maps);
    perm_free(&p_lo); perm_free(&q_lo); perm_free(&p_hi); perm_free(&q_hi);
}

static perm_t perm_from_r2c(const int32_t *r2c, int32_t size)
{
    perm_t p = perm_alloc(size);
    for (int32_t r = 0; r < size; ++r)
        if (r2c[r] >= 0 && r2c[r] < size) perm_set(&p, r, r2c[r]);
    return p;
}

void orc_steady_ant(const int32_t *p_r2c, const int32_t *q_r2c, int32_t k, int32_t *out_r2c)
{
    perm_t p = perm_from_r2c(p_r2c, k), q = perm_from_r2c(q_r2c, k);
    steady_ant_rec(&p, &q);
    memcpy(out_r2c, p.r2c, sizeof(int32_t) * (size_t)k);
    perm_free(&p); perm_free(&q);
}

/* braid_multiplication.hpp:273-406 (staggered_multiply<RowGlue>). */
int32_t orc_staggered_multiply(const int32_t *p_r2c, int32_t p_size,
                               const int32_t *q_r2c, int32_t q_size,
                               int32_t k, int32_t row_glue,
                               int32_t *out_r2c, int32_t *out_c2r)
{
    perm_t p = perm_from_r2c(p_r2c, p_size), q = perm_from_r2c(q_r2c, q_size);
    int32_t size = p_size + q_size - k;
    perm_t prod = perm_alloc(size);
    if (k == 0) {
        for (int32_t i = 0; i < p.size; ++i) {
            int32_t c = p.r2c[i];
            if (c == -1) continue;
            if (row_glue) perm_set(&prod, i + q.size, c + q.size);
            else          perm_set(&prod, i, c);
        }
        for (int32_t i = 0; i < q.size; ++i) {
            int32_t c = q.r2c[i];
            if (c == -1) continue;
            if (row_glue) perm_set(&prod, i, c);
            else          perm_set(&prod, i + p.size, c + p.size);
        }
    } else {
        perm_t p_red = perm_alloc(k), q_red = perm_alloc(k);
        int32_t *map_row = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
        int32_t *map_col = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
        if (row_glue) {
            vert_slice(&p, 0, k, map_row, &p_red);
            horz_slice(&q, q.size - k, q.size, map_col, &q_red);
        } else {
            vert_slice(&p, p.size - k, p.size, map_row, &p_red);
            horz_slice(&q, 0, k, map_col, &q_red);
        }
        steady_ant_rec(&p_red, &q_red);
        for (int32_t i = 0; i < k; ++i) {
            int32_t cur_col = map_col[p_red.r2c[i]], cur_row = map_row[i];
            if (row_glue) perm_set(&prod, q.size - k + cur_row, cur_col);
            else          perm_set(&prod, cur_row, cur_col + p.size - k);
        }
        if (row_glue) {
            for (int32_t j = k; j < p.size; ++j) {
                int32_t r = p.c2r[j];
                if (r != -1) perm_set(&prod, r + q.size - k, j + q.size - k);
            }
            for (int32_t i = 0; i < q.size - k; ++i) {
                int32_t c = q.r2c[i];
                if (c != -1) perm_set(&prod, i, c);
            }
        } else {
            for (int32_t j = 0; j < p.size - k; ++j) {
                int32_t r = p.c2r[j];
                if (r != -1) perm_set(&prod, r, j);
            }
            for (int32_t i = k; i < q.size; ++i) {
                int32_t c = q.r2c[i];
                if (c != -1) perm_set(&prod, i - k + p.size, c - k + p.size);
            }
        }
        free(map_row); free(map_col);
        perm_free(&p_red); perm_free(&q_red);
    }
    memcpy(out_r2c, prod.r2c, sizeof(int32_t) * (size_t)size);
    if (out_c2r) memcpy(out_c2r, prod.c2r, sizeof(int32_t) * (size_t)size);
    perm_free(&p); perm_free(&q); perm_free(&prod);
    return size;
}

/* Distance product from its definition (Tiskin's unit-Monge product). */
void orc_sticky_bruteforce(const int32_t *p_r2c, const int32_t *q_r2c, int32_t k, int32_t *out_r2c)
{
    int32_t W = k + 1;
    int32_t *PS = (int32_t *)calloc((size_t)W * W, sizeof(int32_t));
    int32_t *QS = (int32_t *)calloc((size_t)W * W, sizeof(int32_t));
    int32_t *RS = (int32_t *)calloc((size_t)W * W, sizeof(int32_t));
    /* X^S(i,j) = #{ (r,c): r >= i, c < j } */
    for (int32_t i = 0; i <= k; ++i)
        for (int32_t j = 0; j <= k; ++j) {
            int32_t sp = 0, sq = 0;
            for (int32_t r = i; r < k; ++r) { sp += p_r2c[r] < j; sq += q_r2c[r] < j; }
            PS[i * W + j] = sp; QS[i * W + j] = sq;
        }
    for (int32_t i = 0; i <= k; ++i)
        for (int32_t l = 0; l <= k; ++l) {
            int32_t best = 1 << 30;
            for (int32_t j = 0; j <= k; ++j) {
                int32_t s = PS[i * W + j] + QS[j * W + l];
                if (s < best) best = s;
            }
            RS[i * W + l] = best;
        }
    for (int32_t i = 0; i < k; ++i) {
        out_r2c[i] = -1;
        for (int32_t c = 0; c < k; ++c) {
            int32_t d = RS[i * W + c + 1] - RS[i * W + c] - RS[(i + 1) * W + c + 1] + RS[(i + 1) * W + c];
            if (d == 1) out_r2c[i] = c;
        }
    }
    free(PS); free(QS); free(RS);
}
