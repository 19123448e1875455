// Single-workgroup FP64 tile factorizations held in LDS (tiles up to 64x64).
//
//   potrf : L = chol(A) (lower, strict upper zero-filled), like
//           scipy.linalg.cholesky(lower=True) which the reference uses
//           (src/serinv/__init__.py:151).  A non-positive or NaN pivot at
//           tile column j records info = row_offset + j + 1 (LAPACK style).
//   trtri : X = L^{-1} for a lower-triangular L (strict upper zero-filled).
//
// Tiles are described by a batch of (pointer, ld) computed from two batch
// strides, so one launch can factor or invert many independent tiles.
#pragma once
#include "common.hpp"

namespace rocbta {

constexpr int TT = 64;  // maximum tile order
constexpr int TT_THREADS = 256;
constexpr int TT_LD = TT + 1;

struct TileBatch {
  double* A;         // tile (0,0) of batch entry 0
  int64_t lda;
  int64_t sA1, sA2;  // batch strides
  double* X;         // optional output (inverse); may alias nothing in A
  int64_t ldx;
  int64_t sX1, sX2;
  int m;             // tile order (<= 64)
  int batch1, batch2;
  int* info;
  int64_t info_row0;       // global row of the tile's first row, for info
  int64_t info_s1, info_s2;  // global-row strides along the batch dims
  // Strict-upper zero fill of the rows of the tile that lie right of it:
  // zcols columns starting at column m of each tile row (0 = none).
  int zcols;
};

// Cholesky of S (m x m, lower part valid) in place; all threads participate.
__device__ void lds_potrf(double (*S)[TT_LD], int m, int* info, int64_t row0) {
  const int tid = threadIdx.x;
  for (int j = 0; j < m; ++j) {
    if (tid == 0) {
      double d = S[j][j];
      if (!(d > 0.0)) {  // also catches NaN
        record_info(info, (int)(row0 + j + 1));
        d = __builtin_nan("");
      }
      S[j][j] = sqrt(d);
    }
    __syncthreads();
    const double inv = 1.0 / S[j][j];
    if (tid > j && tid < m) S[tid][j] *= inv;
    __syncthreads();
    // trailing update of the lower triangle: columns k in (j, m)
    const int k = tid & 63, g = tid >> 6;
    if (k > j && k < m) {
      const double lkj = S[k][j];
      for (int i = k + g; i < m; i += 4) S[i][k] -= S[i][j] * lkj;
    }
    __syncthreads();
  }
}

// X = L^{-1}, L lower (m x m) in S; X written to Xs (lower, upper zero).
// Four lanes per column j: lane q sums the k = q (mod 4) part of the
// forward substitution and keeps the X[k][j] it produced; the partial sums
// are combined with lane shuffles, so no cross-lane LDS traffic is needed.
__device__ void lds_trtri(const double (*S)[TT_LD], double (*Xs)[TT_LD], int m) {
  const int tid = threadIdx.x;
  const int j = tid >> 2, q = tid & 3;
  for (int i = 0; i < m; ++i) {
    if (j < m && i >= j) {
      double s = 0.0;
      for (int k = j + ((q - j) & 3); k < i; k += 4) s += S[i][k] * Xs[k][j];
      s += __shfl_xor(s, 1, 4);
      s += __shfl_xor(s, 2, 4);
      const double x = ((i == j ? 1.0 : 0.0) - s) / S[i][i];
      if ((i & 3) == q) Xs[i][j] = x;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < TT * TT; idx += TT_THREADS) {
    int r = idx / TT, c = idx % TT;
    if (r < m && c < m && c > r) Xs[r][c] = 0.0;
  }
  __syncthreads();
}

// mode 0: potrf (+ inverse into X if X != nullptr); mode 1: trtri only.
template <int MODE>
__global__ void __launch_bounds__(TT_THREADS) tile_factor_kernel(TileBatch tb) {
  __shared__ double S[TT][TT_LD];
  __shared__ double Xs[TT][TT_LD];
  const int bz = blockIdx.x;
  const int b1 = bz % tb.batch1, b2 = bz / tb.batch1;
  double* A = tb.A + b1 * tb.sA1 + b2 * tb.sA2;
  const int m = tb.m;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < TT * TT; idx += TT_THREADS) {
    int r = idx / TT, c = idx % TT;
    double v = 0.0;
    if (r < m && c <= r) v = A[(int64_t)r * tb.lda + c];
    S[r][c] = v;
  }
  __syncthreads();
  if (MODE == 0) {
    lds_potrf(S, m, tb.info, tb.info_row0 + b1 * tb.info_s1 + b2 * tb.info_s2);
    // write L back, upper triangle zero-filled
    for (int idx = tid; idx < m * m; idx += TT_THREADS) {
      int r = idx / m, c = idx % m;
      A[(int64_t)r * tb.lda + c] = (c <= r) ? S[r][c] : 0.0;
    }
    if (tb.zcols > 0) {
      for (int r = 0; r < m; ++r)
        for (int c = tid; c < tb.zcols; c += TT_THREADS) A[(int64_t)r * tb.lda + m + c] = 0.0;
    }
  }
  if (tb.X != nullptr) {
    lds_trtri(S, Xs, m);
    double* X = tb.X + b1 * tb.sX1 + b2 * tb.sX2;
    for (int idx = tid; idx < m * m; idx += TT_THREADS) {
      int r = idx / m, c = idx % m;
      X[(int64_t)r * tb.ldx + c] = Xs[r][c];
    }
  }
}

}  // namespace rocbta
