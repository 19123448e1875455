// Small FP64 kernels for the BTA solve and log-determinant.
#pragma once
#include "common.hpp"

namespace rocbta {

constexpr int SV_COLS = 16;  // right-hand-side columns per workgroup (4 per wave)
constexpr int SV_THREADS = 256;

__device__ __forceinline__ double sv_readlane(double v, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Solve op(L) X = B in place for one lower-triangular tile L (m x m, m <= 64)
// and the m x r block B (row stride ldb); op = identity (trans = 0) or
// transpose (trans = 1).  Lane i of every wave owns row i; wave w of
// workgroup x owns columns 16x + 4w .. 16x + 4w + 3.  Each substitution step
// broadcasts the solved row by readlane, so the 64 sequential steps need no
// workgroup barrier.
__global__ void __launch_bounds__(SV_THREADS)
    trsv_tile_kernel(const double* Lt, int64_t ld, int m, double* B, int64_t ldb, int r,
                     int trans) {
  __shared__ double S[64][65];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < 64 * 64; idx += SV_THREADS) {
    int i = idx / 64, k = idx % 64;
    S[i][k] = (i < m && k <= i) ? Lt[(int64_t)i * ld + k] : 0.0;
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  const int c0 = blockIdx.x * SV_COLS + wave * 4;
  if (c0 >= r) return;  // whole wave idle (no later barrier)
  const bool row_ok = lane < m;
  const double invd = row_ok ? 1.0 / S[lane][lane] : 0.0;
  double x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    x[q] = (row_ok && c0 + q < r) ? B[(int64_t)lane * ldb + c0 + q] : 0.0;
  if (!trans) {
    for (int j = 0; j < m; ++j) {
      const double lij = (lane > j && row_ok) ? S[lane][j] : 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (lane == j) x[q] *= invd;
        const double xj = sv_readlane(x[q], j);
        x[q] -= lij * xj;
      }
    }
  } else {
    for (int j = m - 1; j >= 0; --j) {
      const double lji = (lane < j) ? S[j][lane] : 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (lane == j) x[q] *= invd;
        const double xj = sv_readlane(x[q], j);
        x[q] -= lji * xj;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (row_ok && c0 + q < r) B[(int64_t)lane * ldb + c0 + q] = x[q];
}

// out = 2 * (sum_i sum_j log D[i][j][j] + sum_j log T[j][j]); deterministic
// single-workgroup reduction.
__device__ __forceinline__ void logdet_body(const double* D, int64_t n, int64_t b,
                                            const double* T, int64_t a, double* out,
                                            double* red) {
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

// Multi-workgroup form: workgroup w sums log diag over the rows
// [w * rows_per_wg, (w + 1) * rows_per_wg) of the N = n * b + a diagonal
// entries (block rows first, then the tip) into part[w]; logdet_finish_kernel
// then adds the partials in index order, so the result does not depend on
// scheduling.  Each thread handles consecutive rows: the diagonal entries
// of one block are b + 1 doubles apart.
constexpr int LD_ROWS_PER_WG = 1024;

__device__ __forceinline__ double diag_entry(const double* D, int64_t n, int64_t b,
                                             const double* T, int64_t a, int64_t row) {
  if (row < n * b) {
    const int64_t i = row / b, j = row - i * b;
    return D[i * b * b + j * (b + 1)];
  }
  const int64_t j = row - n * b;
  return T[j * (a + 1)];
}

__global__ void __launch_bounds__(256)
    logdet_part_kernel(const double* D, int64_t n, int64_t b, const double* T, int64_t a,
                       double* part) {
  __shared__ double red[256];
  const int64_t N = n * b + a;
  const int64_t r0 = (int64_t)blockIdx.x * LD_ROWS_PER_WG;
  const int64_t r1 = r0 + LD_ROWS_PER_WG < N ? r0 + LD_ROWS_PER_WG : N;
  double s = 0.0;
  for (int64_t row = r0 + threadIdx.x; row < r1; row += 256) s += log(diag_entry(D, n, b, T, a, row));
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) logdet_finish_kernel(const double* part, int nparts,
                                                            double* out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < nparts; k += 256) s += part[k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = 2.0 * red[0];
}

// One workgroup per matrix (batched plans).
struct LogdetDesc {
  const double* D;
  const double* T;
  int64_t n, b, a;
  double* out;
};
__global__ void __launch_bounds__(256) logdet_multi_kernel(const LogdetDesc* d) {
  __shared__ double red[256];
  const LogdetDesc& q = d[blockIdx.x];
  logdet_body(q.D, q.n, q.b, q.T, q.a, q.out, red);
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
