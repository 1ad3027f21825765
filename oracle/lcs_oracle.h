/*
 * CPU oracle for the semi-local LCS seaweed hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU and in plain
 * C, the algorithms of the reference (egor-bv/platform-independent-lcs) so
 * that the HIP product path can be checked against it.  Only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg may load it, and only as the
 * checker or the reported CPU baseline; the product never calls it.
 *
 * Conventions (reference src/lcs_reference.hpp:5-69):
 *   - symbols and strand ids are int32; symbols compare with ==, strands with
 *     signed >;
 *   - the grid row i holds a[m-1-i] (a is reversed), column j holds b[j];
 *   - initial strands h[i] = i, v[j] = m + j;
 *   - a cell swaps its two strands iff a_i == b_j or h > v.
 *
 * Every function names the reference lines it follows.  How each one is
 * pinned (golden vectors, brute force, invariants) is listed in the header
 * comment of oracle/lcs_oracle.c.
 */
#ifndef LCS_ORACLE_H
#define LCS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Sequential comb in the reference's anti-diagonal order
 * (lcs_reference.hpp:34-58).  h_out has m entries, v_out n entries. */
void orc_comb_antidiag(const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                       int32_t *h_out, int32_t *v_out);

/* Same result, row-major order (any order respecting the cell dependencies
 * (i,j-1) and (i+1,j) gives the same strands). */
void orc_comb_rows(const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                   int32_t *h_out, int32_t *v_out);

/* Block-wavefront comb over `threads` OpenMP threads (restatement of the
 * work split of lcs_tiled.hpp:95-137 on the CPU).  Same result. */
void orc_comb_mt(const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                 int32_t *h_out, int32_t *v_out, int32_t threads);

/* Vectorised multithreaded comb, the CPU shape of the reference's tiled
 * comb (Lcs_Semi_Tiled_MT, lcs_tiled.hpp:95-137, with update_stripe,
 * lcs_common.hpp:131-173): the W lanes of a SIMD register own W*TM rows of a
 * stripe and sweep its columns along skewed anti-diagonals; stripes and
 * column blocks run as a pipeline over `threads` threads.  isa: 0 = best the
 * host offers, 1 = AVX2 (W = 8), 2 = AVX-512 (W = 16), 3 = scalar
 * (orc_comb_mt).  Returns the isa used, or -1 if the host lacks the one
 * asked for.  Same result as orc_comb_rows. */
int32_t orc_comb_simd(const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                      int32_t *h_out, int32_t *v_out, int32_t threads, int32_t isa);

/* Comb a sub-rectangle given prepared (reversed) a, b and current strands,
 * the reference's residual fixup (lcs_residual_fixup.hpp:74-122). */
void orc_fixup(const int32_t *a_prepared, const int32_t *b_prepared,
               int32_t *h_strands, int32_t *v_strands,
               int32_t i0, int32_t isize, int32_t j0, int32_t jsize);

/* PermutationMatrix::FromStrands (permutation_matrix.hpp:88-115).
 * row_to_col / col_to_row have m+n entries, -1 where unset. */
void orc_from_strands(const int32_t *h, int32_t m, const int32_t *v, int32_t n,
                      int32_t *row_to_col, int32_t *col_to_row);

/* PermutationMatrix::hash (permutation_matrix.hpp:192-203), restated without
 * signed overflow: unsigned 64-bit arithmetic, reinterpreted as int64 before
 * the C remainder, which reproduces the two's-complement wrap. */
int64_t orc_hash(const int32_t *col_to_row, int32_t size);

/* Global LCS read from the strands: #{ j : v_out[j] < m }. */
int32_t orc_llcs_from_strands(const int32_t *v_out, int32_t m, int32_t n);

/* String-substring semi-local scores H(i,j) = LCS(a, b[i:j)) for all
 * 0 <= i <= j <= n, read from the permutation row_to_col (size m+n).
 * out is (n+1)*(n+1), row-major, entries with i > j set to 0. */
void orc_string_substring(const int32_t *row_to_col, int32_t m, int32_t n,
                          int32_t *out);

/* Textbook O(mn) LCS length (brute force, independent of the comb). */
int32_t orc_llcs_dp(const int32_t *a, int32_t m, const int32_t *b, int32_t n);

/* LCS length by the word-parallel recurrence over 64-bit words (the
 * algorithm the reference's disabled Lcs_Prefix_Binary_Reference,
 * lcs_reference.hpp:135-211, packs into words; Hyyro's form of the
 * Allison-Dix recurrence): V = all ones over m bits; per symbol c of b,
 * U = V & M[c], V = (V + U) | (V & ~M[c]); LCS = zero bits of V.  Any int32
 * symbols (the masks are built per distinct symbol of a).  threads > 1
 * splits nothing (the recurrence is sequential in b); it exists for the
 * timing harness only and must be 1. */
int32_t orc_llcs_bitpar(const int32_t *a, int32_t m, const int32_t *b, int32_t n);

/* Semi-local quadrant queries by direct count.  With the seaweed matrix
 * H(i, j) = (j - i) - #{ r >= m + i : P[r] < j }, i in [-m, n], j in [0, m+n]
 * (b padded with m wildcards on each side), kind:
 *   0 LCS(a, b[q0:q1))    = H(q0, q1)
 *   1 LCS(a[q0:q1), b)    = H(-q0, n + m - q1) - q0 - (m - q1)
 *   2 LCS(a[:q0), b[q1:)) = H(q1, n + m - q0) - (m - q0)
 *   3 LCS(a[q0:), b[:q1)) = H(-q0, q1) - q0
 *   4 #{ r >= q0 : P[r] < q1 }
 * -1 for a query outside its quadrant.  The reference reads no scores from
 * its permutations; the read-out is the semi-local LCS definition. */
void orc_semilocal(const int32_t *row_to_col, int32_t m, int32_t n, int32_t kind,
                   const int32_t *q0, const int32_t *q1, int32_t nq, int32_t *out);

/* staggered_multiply<RowGlue>(p, q, k) of braid_multiplication.hpp:273-406
 * with its steady ant (:5-270).  p, q are given by row_to_col; the product
 * has p_size + q_size - k points.  out_c2r may be NULL. */
int32_t orc_staggered_multiply(const int32_t *p_r2c, int32_t p_size,
                               const int32_t *q_r2c, int32_t q_size,
                               int32_t k, int32_t row_glue,
                               int32_t *out_r2c, int32_t *out_c2r);

/* The sticky (unit-Monge distance) product of two k-permutations computed
 * from its definition: R^S(i,l) = min_j P^S(i,j) + Q^S(j,l), with
 * X^S(i,j) = #{ (r,c) in X : r >= i, c < j }.  O(k^3); tiny k only. */
void orc_sticky_bruteforce(const int32_t *p_r2c, const int32_t *q_r2c,
                           int32_t k, int32_t *out_r2c);

/* The steady-ant core alone: sticky product of two k-permutations. */
void orc_steady_ant(const int32_t *p_r2c, const int32_t *q_r2c, int32_t k,
                    int32_t *out_r2c);

#ifdef __cplusplus
}
#endif
#endif
