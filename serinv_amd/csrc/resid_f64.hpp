// FP64 residual Out = B - A X of an SPD BTA matrix A given by its stored
// blocks (the lower triangles of D and T, L = A_{i+1,i}, R = A_{t,i}), for
// the iterative refinement of solves with a mixed-precision factor (an
// extension; no counterpart in the reference).
//
// Work split (64-row tiles, RC right-hand-side columns per chunk):
//   residual_rows_kernel  grid (NT + n, chunks): workgroup g < NT forms the
//     main row tile g from the row band of A (block i: sym(D_i), A_{i,i-1},
//     A_{i+1,i}^T, A_{t,i}^T); workgroup NT + i forms the arrow-row partial
//     R_i x_i of block i;
//   residual_tip_kernel   grid (ta, chunks): sums the arrow partials in block
//     order and adds sym(T) x_t.
// Both are HBM-bound GEMV-type passes (A is read about twice: each lower
// block once as itself and once transposed for the row above).
#pragma once
#include "solve_sweep.hpp"

namespace rocbta {

struct ResidP {
  const double* D;
  const double* L;
  const double* R;
  const double* T;
  const double* X;
  const double* B;
  double* Out;
  int64_t ld;  // row stride of X, B and Out
  int n, b, a, nt, ta, NT;
  int r;
  double* part;  // [chunk][n][ta * 64 * RC]
};

template <int RC>
__device__ __forceinline__ void rs_vec(const ResidP& P, int chunk, int64_t row0, int rows,
                                       double* V) {
  const int cols = min(RC, P.r - chunk * RC);
  for (int idx = threadIdx.x; idx < SW_T * RC; idx += SW_THREADS) {
    const int k = idx / RC, c = idx % RC;
    V[idx] = (k < rows && c < cols) ? P.X[(row0 + k) * P.ld + chunk * RC + c] : 0.0;
  }
}

// acc += op(tile) x for one 64 x 64 tile (rows x cols stored at p, stride ld).
// mode 0: tile as stored; 1: transposed; 2: symmetric from the stored lower
// triangle (diagonal tiles of D and T).
template <int RC>
__device__ __forceinline__ void rs_tile(const ResidP& P, const double* p, int64_t ld, int rows,
                                        int cols, int mode, int chunk, int64_t xrow0, int xrows,
                                        double* A, double* V, double (&acc)[RC]) {
  SwRegs x;
  sw_load<false>(p, ld, rows, cols, x);
  __syncthreads();  // previous product done with A / V
  sw_store_lds(x, A);
  rs_vec<RC>(P, chunk, xrow0, xrows, V);
  __syncthreads();
  if (mode == 2) {
    for (int idx = threadIdx.x; idx < SW_T * SW_T; idx += SW_THREADS) {
      const int rr = idx >> 6, cc = idx & 63;
      if (cc > rr) A[rr * SW_LD + cc] = A[cc * SW_LD + rr];
    }
    __syncthreads();
  }
  if (mode == 1)
    sw_gemv<RC, true, false>(A, V, acc);
  else
    sw_gemv<RC, false, false>(A, V, acc);
}

template <int RC>
struct RsLds {
  double A[SW_T * SW_LD];
  double V[SW_T * RC];
  double red[4 * SW_T * RC];
  double O[SW_T * RC];
};

template <int RC>
__global__ void __launch_bounds__(SW_THREADS) residual_rows_kernel(ResidP P) {
  __shared__ __attribute__((aligned(16))) RsLds<RC> S;
  const int chunk = blockIdx.y;
  const int cols = min(RC, P.r - chunk * RC);
  const int64_t b = P.b, bb = b * b;
  double acc[RC];
#pragma unroll
  for (int c = 0; c < RC; ++c) acc[c] = 0.0;
  if ((int)blockIdx.x >= P.NT) {  // arrow partial R_i x_i of block i
    const int i = blockIdx.x - P.NT;
    for (int t = 0; t < P.ta; ++t) {
      const int mt = min(SW_T, P.a - t * SW_T);
#pragma unroll
      for (int c = 0; c < RC; ++c) acc[c] = 0.0;
      for (int cc = 0; cc < P.nt; ++cc) {
        const int mc = min(SW_T, P.b - cc * SW_T);
        rs_tile<RC>(P, P.R + (int64_t)i * P.a * b + (int64_t)t * SW_T * b + cc * SW_T, b, mt, mc, 0,
                    chunk, i * b + cc * SW_T, mc, S.A, S.V, acc);
      }
      sw_reduce<RC>(acc, S.red, S.O);
      double* pp = P.part + (((int64_t)chunk * P.n + i) * P.ta + t) * SW_T * RC;
      for (int idx = threadIdx.x; idx < SW_T * RC; idx += SW_THREADS) pp[idx] = S.O[idx];
    }
    return;
  }
  const int g = blockIdx.x;
  const int i = g / P.nt, q = g % P.nt;
  const int mq = min(SW_T, P.b - q * SW_T);
  const double* Di = P.D + (int64_t)i * bb;
  for (int cc = 0; cc < P.nt; ++cc) {  // sym(D_i)
    const int mc = min(SW_T, P.b - cc * SW_T);
    if (cc < q)
      rs_tile<RC>(P, Di + (int64_t)q * SW_T * b + cc * SW_T, b, mq, mc, 0, chunk, i * b + cc * SW_T,
                  mc, S.A, S.V, acc);
    else if (cc > q)
      rs_tile<RC>(P, Di + (int64_t)cc * SW_T * b + q * SW_T, b, mc, mq, 1, chunk, i * b + cc * SW_T,
                  mc, S.A, S.V, acc);
    else
      rs_tile<RC>(P, Di + (int64_t)q * SW_T * (b + 1), b, mq, mq, 2, chunk, i * b + cc * SW_T, mc,
                  S.A, S.V, acc);
  }
  if (i > 0) {  // A_{i,i-1} = L[i-1]
    const double* Lm = P.L + (int64_t)(i - 1) * bb;
    for (int cc = 0; cc < P.nt; ++cc) {
      const int mc = min(SW_T, P.b - cc * SW_T);
      rs_tile<RC>(P, Lm + (int64_t)q * SW_T * b + cc * SW_T, b, mq, mc, 0, chunk,
                  (int64_t)(i - 1) * b + cc * SW_T, mc, S.A, S.V, acc);
    }
  }
  if (i + 1 < P.n) {  // A_{i,i+1} = L[i]^T
    const double* Lp = P.L + (int64_t)i * bb;
    for (int cc = 0; cc < P.nt; ++cc) {
      const int mc = min(SW_T, P.b - cc * SW_T);
      rs_tile<RC>(P, Lp + (int64_t)cc * SW_T * b + q * SW_T, b, mc, mq, 1, chunk,
                  (int64_t)(i + 1) * b + cc * SW_T, mc, S.A, S.V, acc);
    }
  }
  for (int t = 0; t < P.ta; ++t) {  // A_{i,t} = R_i^T
    const int mt = min(SW_T, P.a - t * SW_T);
    rs_tile<RC>(P, P.R + (int64_t)i * P.a * b + (int64_t)t * SW_T * b + q * SW_T, b, mt, mq, 1,
                chunk, (int64_t)P.n * b + t * SW_T, mt, S.A, S.V, acc);
  }
  sw_reduce<RC>(acc, S.red, S.O);
  const int64_t row0 = (int64_t)i * b + q * SW_T;
  for (int idx = threadIdx.x; idx < mq * RC; idx += SW_THREADS) {
    const int k = idx / RC, c = idx % RC;
    if (c < cols) {
      const int64_t o = (row0 + k) * P.ld + chunk * RC + c;
      P.Out[o] = P.B[o] - S.O[idx];
    }
  }
}

template <int RC>
__global__ void __launch_bounds__(SW_THREADS) residual_tip_kernel(ResidP P) {
  __shared__ __attribute__((aligned(16))) RsLds<RC> S;
  const int t = blockIdx.x, chunk = blockIdx.y;
  const int cols = min(RC, P.r - chunk * RC);
  const int o = threadIdx.x & 63, kq = threadIdx.x >> 6;
  const int mt = min(SW_T, P.a - t * SW_T);
  double acc[RC];
#pragma unroll
  for (int c = 0; c < RC; ++c) acc[c] = 0.0;
  for (int i = kq; i < P.n; i += 4) {
    const double* pp = P.part + (((int64_t)chunk * P.n + i) * P.ta + t) * SW_T * RC + o * RC;
#pragma unroll
    for (int c = 0; c < RC; ++c) acc[c] += pp[c];
  }
  const int64_t tip0 = (int64_t)P.n * P.b;
  for (int u = 0; u < P.ta; ++u) {  // sym(T)
    const int mu = min(SW_T, P.a - u * SW_T);
    if (u < t)
      rs_tile<RC>(P, P.T + (int64_t)t * SW_T * P.a + u * SW_T, P.a, mt, mu, 0, chunk,
                  tip0 + u * SW_T, mu, S.A, S.V, acc);
    else if (u > t)
      rs_tile<RC>(P, P.T + (int64_t)u * SW_T * P.a + t * SW_T, P.a, mu, mt, 1, chunk,
                  tip0 + u * SW_T, mu, S.A, S.V, acc);
    else
      rs_tile<RC>(P, P.T + (int64_t)t * SW_T * (P.a + 1), P.a, mt, mt, 2, chunk, tip0 + u * SW_T,
                  mu, S.A, S.V, acc);
  }
  sw_reduce<RC>(acc, S.red, S.O);
  for (int idx = threadIdx.x; idx < mt * RC; idx += SW_THREADS) {
    const int k = idx / RC, c = idx % RC;
    if (c < cols) {
      const int64_t off = (tip0 + t * SW_T + k) * P.ld + chunk * RC + c;
      P.Out[off] = P.B[off] - S.O[idx];
    }
  }
}

}  // namespace rocbta
