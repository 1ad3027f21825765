// Semi-local queries on a seaweed permutation: the quadrant -> dominance map
// shared by host validation and the device kernels, and the wavelet-matrix
// index layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lcs_hip.h"

namespace lcsq {

// Query q = (q0, q1) of `kind` on P(a, b) (m rows of a, n columns of b) as
//     result = sign * (base - D(r0, c0)) + (1 - sign) * D(r0, c0),
//     D(r0, c0) = #{ r >= r0 : P[r] < c0 }.
// The four quadrants are H(i, j) = (j - i) - D(m + i, j) at
//   string-substring  (i, j) = (q0, q1)               H
//   substring-string  (i, j) = (-q0, n + m - q1)      H - q0 - (m - q1) = n - D
//   prefix-suffix     (i, j) = (q1, n + m - q0)       H - (m - q0)       = (n - q1) - D
//   suffix-prefix     (i, j) = (-q0, q1)              H - q0             = q1 - D
// Returns false when the query lies outside its quadrant.
struct Window {
    int64_t r0, c0;
    int32_t base;
    bool score;   // false: the raw dominance count
};

__host__ __device__ inline bool query_window(int kind, int32_t m, int32_t n, int32_t q0, int32_t q1, Window &w)
{
    const int64_t M = m, N = n;
    switch (kind) {
    case LCS_SEMI_STRING_SUBSTRING:
        if (q0 < 0 || q0 > q1 || q1 > n) return false;
        w = {M + q0, q1, q1 - q0, true};
        return true;
    case LCS_SEMI_SUBSTRING_STRING:
        if (q0 < 0 || q0 > q1 || q1 > m) return false;
        w = {M - q0, N + M - q1, n, true};
        return true;
    case LCS_SEMI_PREFIX_SUFFIX:
        if (q0 < 0 || q0 > m || q1 < 0 || q1 > n) return false;
        w = {M + q1, N + M - q0, n - q1, true};
        return true;
    case LCS_SEMI_SUFFIX_PREFIX:
        if (q0 < 0 || q0 > m || q1 < 0 || q1 > n) return false;
        w = {M - q0, q1, q1, true};
        return true;
    case LCS_SEMI_DOMINANCE:
        w = {q0 < 0 ? 0 : q0, q1, 0, false};
        return true;
    }
    return false;
}

// One level of the wavelet matrix: 64 rows per word.  `bits` holds bit
// (L-1-level) of the level's values, `zeros_before` the number of zero bits
// in all earlier words of the level.  16 bytes, so a rank is one load.
struct RankWord {
    unsigned long long bits;
    uint32_t zeros_before;
    uint32_t pad;
};

}  // namespace lcsq

// Launchers (query_kernels.hip); 0 or a negative LCS_E* status.
int lcsq_launch_scan(const int32_t *perm, int32_t size, int kind, int32_t m, int32_t n, const int32_t *q0,
                     const int32_t *q1, int32_t nq, int32_t *out, hipStream_t s);
// Builds the index of `perm` (size rows) into words[levels * W] and
// zeros[levels], W = size/64 + 1, levels = bit length of size.  scratch holds
// 2*size + blocks + 1 int32 (blocks = ceil(W/64)); scratch[last] becomes 1 if
// a value lies outside [-1, size).
int lcsq_launch_index_build(const int32_t *perm, int32_t size, int levels, lcsq::RankWord *words, uint32_t *zeros,
                            int32_t *scratch, hipStream_t s);
int lcsq_launch_index_query(const lcsq::RankWord *words, const uint32_t *zeros, int levels, int32_t size, int kind,
                            int32_t m, int32_t n, const int32_t *q0, const int32_t *q1, int32_t nq, int32_t *out,
                            hipStream_t s);
