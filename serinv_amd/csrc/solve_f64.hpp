// Small FP64 kernels for the BTA solve and log-determinant.
#pragma once
#include "common.hpp"

namespace rocbta {

constexpr int SV_COLS = 64;  // right-hand-side columns per workgroup
constexpr int SV_THREADS = 256;

// Solve op(L) X = B in place for one lower-triangular tile L (m x m, m <= 64)
// and the m x r block B (row stride ldb); op = identity (trans = 0) or
// transpose (trans = 1).  Workgroup x handles columns [64x, 64x + 64).
// Four lanes per column split the substitution sums (k = q mod 4) and combine
// them with lane shuffles; each lane only re-reads solution entries it wrote.
__global__ void __launch_bounds__(SV_THREADS)
    trsv_tile_kernel(const double* Lt, int64_t ld, int m, double* B, int64_t ldb, int r,
                     int trans) {
  __shared__ double S[64][65];
  __shared__ double Xs[64][SV_COLS + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 64 * 64; idx += SV_THREADS) {
    int i = idx / 64, k = idx % 64;
    S[i][k] = (i < m && k <= i) ? Lt[(int64_t)i * ld + k] : 0.0;
  }
  __syncthreads();
  const int c = tid >> 2, q = tid & 3;
  const int col = blockIdx.x * SV_COLS + c;
  const bool active = col < r;
  if (!trans) {
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      if (active)
        for (int k = q; k < i; k += 4) s += S[i][k] * Xs[k][c];
      s += __shfl_xor(s, 1, 4);
      s += __shfl_xor(s, 2, 4);
      if (active && (i & 3) == q) Xs[i][c] = (B[(int64_t)i * ldb + col] - s) / S[i][i];
    }
  } else {
    for (int i = m - 1; i >= 0; --i) {
      double s = 0.0;
      if (active)
        for (int k = i + 1 + ((q - i - 1) & 3); k < m; k += 4) s += S[k][i] * Xs[k][c];
      s += __shfl_xor(s, 1, 4);
      s += __shfl_xor(s, 2, 4);
      if (active && (i & 3) == q) Xs[i][c] = (B[(int64_t)i * ldb + col] - s) / S[i][i];
    }
  }
  // each lane writes back the rows it owns
  if (active)
    for (int i = q; i < m; i += 4) B[(int64_t)i * ldb + col] = Xs[i][c];
}

// out = 2 * (sum_i sum_j log D[i][j][j] + sum_j log T[j][j]); deterministic
// single-workgroup reduction.
__global__ void __launch_bounds__(256)
    logdet_kernel(const double* D, int64_t n, int64_t b, const double* T, int64_t a,
                  double* out) {
  __shared__ double red[256];
  double s = 0.0;
  const int64_t total = n * b;
  for (int64_t idx = threadIdx.x; idx < total; idx += 256) {
    int64_t i = idx / b, j = idx % b;
    s += log(D[i * b * b + j * b + j]);
  }
  for (int64_t j = threadIdx.x; j < a; j += 256) s += log(T[j * a + j]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = 2.0 * red[0];
}

__global__ void fill_f64_kernel(double* p, int64_t count, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

__global__ void set_int_kernel(int* p, int v) {
  if (threadIdx.x == 0) *p = v;
}

}  // namespace rocbta
