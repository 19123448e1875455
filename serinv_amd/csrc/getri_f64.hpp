// General (non-symmetric) inverse of one m x m tile (m <= 64) by in-place
// Gauss-Jordan elimination without pivoting, for the diagonally dominant
// blocks of the non-SPD (ddbt*) routines, where elimination without row
// exchanges is stable.  One workgroup of 256 threads; the tile lives in LDS.
// Thread t owns column t % 64 and rows t / 64 + 4 q (q = 0 .. 15).  Step k:
//   p = 1 / S[k][k];  row k: S[k][j] *= p (j != k), S[k][k] = p;
//   rows i != k: S[i][j] -= S[i][k] S[k][j] (j != k), S[i][k] = -S[i][k] p.
// Padding rows / columns (m < 64) hold the identity and are untouched.
#pragma once
#include "common.hpp"

namespace rocbta {

__global__ void __launch_bounds__(256) tile_getri_kernel(const double* A, int64_t lda, int m,
                                                        double* P) {
  __shared__ double S[64][65];
  const int tid = threadIdx.x;
  const int j = tid & 63, r0 = tid >> 6;
  for (int idx = tid; idx < 64 * 64; idx += 256) {
    const int r = idx >> 6, c = idx & 63;
    S[r][c] = (r < m && c < m) ? A[(int64_t)r * lda + c] : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    const double p = 1.0 / S[k][k];
    const double rk = (j == k) ? p : S[k][j] * p;  // new row k, column j
    double f[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) f[q] = S[r0 + 4 * q][k];
    __syncthreads();  // every read of step k precedes its writes
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int i = r0 + 4 * q;
      if (i == k)
        S[i][j] = rk;
      else if (j == k)
        S[i][j] = -f[q] * p;
      else
        S[i][j] -= f[q] * rk;
    }
    __syncthreads();
  }
  for (int idx = tid; idx < 64 * 64; idx += 256) P[idx] = S[idx >> 6][idx & 63];
}

}  // namespace rocbta
