// General (non-symmetric) inverse of one m x m tile (m <= 64) by in-place
// Gauss-Jordan elimination without pivoting, for the diagonally dominant
// blocks of the non-SPD (ddbt*) routines, where elimination without row
// exchanges is stable.  One workgroup of 256 threads.  Step k:
//   p = 1 / S[k][k];  row k: S[k][j] *= p (j != k), S[k][k] = p;
//   rows i != k: S[i][j] -= S[i][k] S[k][j] (j != k), S[i][k] = -S[i][k] p.
// The tile stays in registers: thread (wave w, lane j) holds S[w + 4q][j],
// q = 0 .. 15.  The wave holding row k scales it and publishes it through a
// double-buffered LDS row (one barrier per step); every wave takes the
// column entries S[i][k] of its rows from lane k by readlane.  Padding rows
// and columns (m < 64) hold the identity, which the steps k < m leave
// unchanged.  A and P may be the same tile: each thread reads its own
// entries before it writes them.
#pragma once
#include "common.hpp"

namespace rocbta {

__device__ __forceinline__ double getri_readlane(double v, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__global__ void __launch_bounds__(256) tile_getri_kernel(const double* A, int64_t lda, int m,
                                                        double* P, int64_t ldp) {
  __shared__ double row[2][64];
  const int tid = threadIdx.x;
  const int j = tid & 63, w = tid >> 6;
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = w + 4 * q;
    v[q] = (i < m && j < m) ? A[(int64_t)i * lda + j] : (i == j ? 1.0 : 0.0);
  }
#pragma unroll
  for (int qk = 0; qk < 16; ++qk) {
#pragma unroll
    for (int wk = 0; wk < 4; ++wk) {
      const int k = 4 * qk + wk;
      if (k < m) {
        if (w == wk) {  // this wave holds row k
          const double piv = getri_readlane(v[qk], k);
          const double pk = 1.0 / piv;
          row[k & 1][j] = (j == k) ? pk : v[qk] * pk;
        }
        __syncthreads();
        const double rk = row[k & 1][j];
        const double pk = row[k & 1][k];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const double f = getri_readlane(v[q], k);  // S[w + 4q][k] before the step
          if (q == qk && w == wk)
            v[q] = rk;
          else if (j == k)
            v[q] = -f * pk;
          else
            v[q] = __builtin_fma(-f, rk, v[q]);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int i = w + 4 * q;
    if (i < m && j < m) P[(int64_t)i * ldp + j] = v[q];
  }
}

}  // namespace rocbta
