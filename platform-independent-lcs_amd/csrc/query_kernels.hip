// Semi-local score queries read from a seaweed permutation.
//
// For the permutation P of a (m) against b (n) in row_to_col form (row =
// strand id, column = end position), the string-substring score is
//     H(i, j) = LCS(a, b[i:j)) = (j - i) - #{ r in [m+i, m+n) : P[r] < j }
// for 0 <= i <= j <= n: the strands entering the top of columns i..n-1 that
// leave through the bottom of a column left of j.  The general query
//     D(r0, c0) = #{ r >= r0 : P[r] < c0 }
// is the dominance count every semi-local quadrant is built from.  One wave
// per query; each lane counts a strided share of the rows.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcsq {

__global__ void dominance_kernel(const int32_t *perm, int32_t size, const int32_t *qr, const int32_t *qc, int32_t nq,
                                 int32_t *out)
{
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= nq) return;
    const int r0 = qr[wave], c0 = qc[wave];
    int cnt = 0;
    for (int r = r0 + lane; r < size; r += 64) cnt += perm[r] < c0;
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    if (lane == 0) out[wave] = cnt;
}

}  // namespace lcsq

int lcsq_launch_dominance(const int32_t *perm, int32_t size, const int32_t *qr, const int32_t *qc, int32_t nq,
                          int32_t *out, hipStream_t s)
{
    if (nq <= 0) return 0;
    const int blocks = (int)(((long long)nq * 64 + 255) / 256);
    hipLaunchKernelGGL(lcsq::dominance_kernel, dim3(blocks), dim3(256), 0, s, perm, size, qr, qc, nq, out);
    return hipGetLastError() == hipSuccess ? 0 : -5;
}
