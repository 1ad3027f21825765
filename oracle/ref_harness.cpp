// Thin C entry points around the reference's own PermutationMatrix
// (src/permutation_matrix.hpp of egor-bv/platform-independent-lcs), compiled
// from the reference source where it lies.  TEST INFRASTRUCTURE: it pins the
// oracle's FromStrands and hash restatements (permutation_matrix.hpp:88-115,
// 192-203).  The output library goes to oracle/_ref/ and is never committed.
#include <cstdint>
#include "permutation_matrix.hpp"

extern "C" int64_t ref_perm_from_strands_hash(int32_t *h, int32_t m, int32_t *v, int32_t n,
                                              int32_t *row_to_col, int32_t *col_to_row)
{
    PermutationMatrix p = PermutationMatrix::FromStrands(h, m, v, n);
    for (int32_t i = 0; i < m + n; ++i) {
        row_to_col[i] = p.get_col(i);
        col_to_row[i] = p.get_row(i);
    }
    return p.hash();
}
