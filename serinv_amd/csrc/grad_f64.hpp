// Gradient assembly for f = gl * log det A + gq * sum_k y_k^T A^{-1} y_k
// with respect to the stored BTA blocks of A (an extension; no counterpart in
// the reference).
//
// With X = A^{-1} (the selected inverse from pobtasi) and x = A^{-1} y:
//   d log det A / dA = X,   d (y^T A^{-1} y) / dA = -x x^T,   d/dy = 2 x.
// A entry stored once in the block arrays stands for one or two entries of A:
//   D_i, T (only the lower triangle is read): weight 2 below the diagonal,
//           1 on it, 0 above;
//   L_i = A_{i+1,i}, R_i = A_{t,i} (their transposes fill the upper part):
//           weight 2.
// So G = w * (gl * X - gq * sum_k x[row] x[col]) entry by entry, where row /
// col are the global indices of the entry.  HBM-bound elementwise pass over
// the pattern (reads X, writes G; the N x r vector stays in L2).
#pragma once
#include "common.hpp"

namespace rocbta {

struct GradP {
  const double* XD;
  const double* XL;
  const double* XR;
  const double* XT;
  const double* x;  // (N, r), row stride ldx
  int64_t ldx;
  int r;
  int64_t n, b, a;
  const double* gl;  // device scalars (may be null = 0)
  const double* gq;
  double* GD;
  double* GL;
  double* GR;
  double* GT;
  double* gy;  // (N, r), row stride ldx (may be null)
  int64_t nD, nL, nR, nT, nY;  // entry counts of each output
};

__device__ __forceinline__ double grad_xx(const GradP& P, int64_t row, int64_t col) {
  double s = 0.0;
  const double* xr = P.x + row * P.ldx;
  const double* xc = P.x + col * P.ldx;
  for (int k = 0; k < P.r; ++k) s = fma(xr[k], xc[k], s);
  return s;
}

__global__ void __launch_bounds__(256) grad_assemble_kernel(GradP P) {
  const double gl = P.gl ? *P.gl : 0.0;
  const double gq = P.gq ? *P.gq : 0.0;
  const int64_t total = P.nD + P.nL + P.nR + P.nT + P.nY;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t bb = P.b * P.b, ab = P.a * P.b;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    int64_t k = e;
    if (k < P.nD) {
      const int64_t i = k / bb, rc = k % bb, r = rc / P.b, c = rc % P.b;
      const double w = (r > c) ? 2.0 : (r == c ? 1.0 : 0.0);
      const double q = (gq != 0.0 && w != 0.0) ? grad_xx(P, i * P.b + r, i * P.b + c) : 0.0;
      P.GD[k] = w * (gl * P.XD[k] - gq * q);
      continue;
    }
    k -= P.nD;
    if (k < P.nL) {
      const int64_t i = k / bb, rc = k % bb, r = rc / P.b, c = rc % P.b;
      const double q = (gq != 0.0) ? grad_xx(P, (i + 1) * P.b + r, i * P.b + c) : 0.0;
      P.GL[k] = 2.0 * (gl * P.XL[k] - gq * q);
      continue;
    }
    k -= P.nL;
    if (k < P.nR) {
      const int64_t i = k / ab, rc = k % ab, r = rc / P.b, c = rc % P.b;
      const double q = (gq != 0.0) ? grad_xx(P, P.n * P.b + r, i * P.b + c) : 0.0;
      P.GR[k] = 2.0 * (gl * P.XR[k] - gq * q);
      continue;
    }
    k -= P.nR;
    if (k < P.nT) {
      const int64_t r = k / P.a, c = k % P.a;
      const double w = (r > c) ? 2.0 : (r == c ? 1.0 : 0.0);
      const double q =
          (gq != 0.0 && w != 0.0) ? grad_xx(P, P.n * P.b + r, P.n * P.b + c) : 0.0;
      P.GT[k] = w * (gl * P.XT[k] - gq * q);
      continue;
    }
    k -= P.nT;
    {
      const int64_t row = k / P.r, c = k % P.r;
      P.gy[row * P.ldx + c] = 2.0 * gq * P.x[row * P.ldx + c];
    }
  }
}

// out = sum_{i < rows, k < r} y[i, k] x[i, k]; deterministic single-workgroup
// reduction.
__global__ void __launch_bounds__(256)
    quad_kernel(const double* y, const double* x, int64_t rows, int r, int64_t ld, double* out) {
  __shared__ double red[256];
  double s = 0.0;
  const int64_t total = rows * r;
  for (int64_t idx = threadIdx.x; idx < total; idx += 256) {
    const int64_t i = idx / r, k = idx % r;
    s = fma(y[i * ld + k], x[i * ld + k], s);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

}  // namespace rocbta
