// Grouped, batched FP64 MFMA GEMM for gfx950.
//
//   C[p] = beta_p * C[p] + sum_t alpha_t * op(A_t) * op(B_t)      (up to 2 terms)
//
// One launch runs up to ROCBTA_MAX_PROBLEMS independent problems (each with a
// 2-D batch); every workgroup owns one 64x64 tile of one C.  Operands are
// row-major with leading dimensions; op() is the identity or the transpose.
// Triangular operands (a stored triangle of zeros) shorten the K loop.
// `uplo` restricts the written part of C to its lower (or upper) triangle,
// which is how the symmetric rank-k (syrk-type) updates are expressed.
//
// Aliasing contract: C may alias A of a term only when the problem has a
// single 64-wide tile column (N <= 64): a workgroup then owns its rows of C
// entirely and reads all of them before it writes.
#pragma once
#include "common.hpp"

namespace rocbta {

constexpr int GT_M = 64;   // tile rows per workgroup
constexpr int GT_N = 64;   // tile cols per workgroup
constexpr int GT_K = 16;   // K slice staged through LDS per step
constexpr int GT_THREADS = 256;
constexpr int GT_LDS_PAD = 4;
#define ROCBTA_MAX_PROBLEMS 4

enum Tri : uint8_t { TRI_NONE = 0, TRI_LOWER = 1, TRI_UPPER = 2 };

struct GemmTerm {
  const double* A;
  const double* B;
  int64_t lda, ldb;
  int64_t sA1, sA2, sB1, sB2;  // batch strides (elements) along batch dims 1 and 2
  int K;
  uint8_t transA, transB;
  uint8_t triA;  // shape of op(A): TRI_LOWER => op(A)(m,k) = 0 for k > m
  uint8_t triB;  // shape of op(B): TRI_LOWER => op(B)(k,n) = 0 for k < n
  double alpha;
};

struct GemmProblem {
  double* C;
  int64_t ldc;
  int64_t sC1, sC2;
  int M, N;
  int batch1, batch2;
  double beta;
  int nterms;
  int uplo;  // TRI_NONE: full C; TRI_LOWER: only row >= col is written
  int tiles_m, tiles_n;
  int tile_begin;  // first global tile index of this problem
  GemmTerm t[2];
};

struct GemmGroup {
  int nprob;
  int total_tiles;
  GemmProblem p[ROCBTA_MAX_PROBLEMS];
};

// Stage one GT_K slice of alpha*op(A) (64 x GT_K) and op(B) (GT_K x 64) into
// registers.  Out-of-range elements are zero.
struct TileRegs {
  double a[4];
  double b[4];
};

__device__ __forceinline__ void gemm_load_slice(const GemmTerm& t, const double* A,
                                                const double* B, int M, int N,
                                                int m0, int n0, int k0, int kend, TileRegs& r) {
  const int tid = threadIdx.x;
  if (!t.transA) {  // A is M x K row-major: 4 consecutive k per thread
    int m = m0 + (tid >> 2), kq = k0 + (tid & 3) * 4;
    const double* src = A + (int64_t)m * t.lda + kq;
#pragma unroll
    for (int j = 0; j < 4; ++j) r.a[j] = (m < M && kq + j < kend) ? t.alpha * src[j] : 0.0;
  } else {  // A stored K x M: 4 consecutive m per thread
    int k = k0 + (tid >> 4), mq = m0 + (tid & 15) * 4;
    const double* src = A + (int64_t)k * t.lda + mq;
#pragma unroll
    for (int j = 0; j < 4; ++j) r.a[j] = (k < kend && mq + j < M) ? t.alpha * src[j] : 0.0;
  }
  if (!t.transB) {  // B is K x N row-major: 4 consecutive n per thread
    int k = k0 + (tid >> 4), nq = n0 + (tid & 15) * 4;
    const double* src = B + (int64_t)k * t.ldb + nq;
#pragma unroll
    for (int j = 0; j < 4; ++j) r.b[j] = (k < kend && nq + j < N) ? src[j] : 0.0;
  } else {  // B stored N x K: 4 consecutive k per thread
    int n = n0 + (tid >> 2), kq = k0 + (tid & 3) * 4;
    const double* src = B + (int64_t)n * t.ldb + kq;
#pragma unroll
    for (int j = 0; j < 4; ++j) r.b[j] = (n < N && kq + j < kend) ? src[j] : 0.0;
  }
}

// LDS holds the slice k-major: As[k][m], Bs[k][n], so that an MFMA operand
// fetch (16 consecutive m or n per k) is a contiguous 128-byte row segment.
__device__ __forceinline__ void gemm_store_slice(const GemmTerm& t, const TileRegs& r,
                                                 double (*As)[GT_M + GT_LDS_PAD],
                                                 double (*Bs)[GT_N + GT_LDS_PAD]) {
  const int tid = threadIdx.x;
  if (!t.transA) {
    int m = tid >> 2, kq = (tid & 3) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) As[kq + j][m] = r.a[j];
  } else {
    int k = tid >> 4, mq = (tid & 15) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) As[k][mq + j] = r.a[j];
  }
  if (!t.transB) {
    int k = tid >> 4, nq = (tid & 15) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) Bs[k][nq + j] = r.b[j];
  } else {
    int n = tid >> 2, kq = (tid & 3) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) Bs[kq + j][n] = r.b[j];
  }
}

// K range of a term restricted by the triangular shape of its operands.
__device__ __forceinline__ void gemm_k_range(const GemmTerm& t, int m0, int n0, int& kbeg,
                                             int& kend) {
  kbeg = 0;
  kend = t.K;
  if (t.triA == TRI_LOWER) kend = min(kend, m0 + GT_M);
  if (t.triA == TRI_UPPER) kbeg = max(kbeg, m0);
  if (t.triB == TRI_LOWER) kbeg = max(kbeg, n0);
  if (t.triB == TRI_UPPER) kend = min(kend, n0 + GT_N);
  kbeg = (kbeg / GT_K) * GT_K;
}

// Accumulate one term into the wave's 2x2 grid of 16x16 MFMA tiles.
__device__ __forceinline__ void gemm_accumulate_term(const GemmTerm& t, const double* A,
                                                     const double* B, int M, int N, int m0,
                                                     int n0, double (*As)[GT_M + GT_LDS_PAD],
                                                     double (*Bs)[GT_N + GT_LDS_PAD],
                                                     f64x4 (&acc)[2][2]) {
  int kbeg, kend;
  gemm_k_range(t, m0, n0, kbeg, kend);
  if (kbeg >= kend) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  TileRegs regs;
  gemm_load_slice(t, A, B, M, N, m0, n0, kbeg, kend, regs);
  for (int k0 = kbeg; k0 < kend; k0 += GT_K) {
    __syncthreads();
    gemm_store_slice(t, regs, As, Bs);
    __syncthreads();
    if (k0 + GT_K < kend) gemm_load_slice(t, A, B, M, N, m0, n0, k0 + GT_K, kend, regs);
#pragma unroll
    for (int kk = 0; kk < GT_K; kk += 4) {
      double a0 = As[kk + lk][wm + li];
      double a1 = As[kk + lk][wm + 16 + li];
      double b0 = Bs[kk + lk][wn + li];
      double b1 = Bs[kk + lk][wn + 16 + li];
      acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
      acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
      acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
      acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
    }
  }
}

__global__ void __launch_bounds__(GT_THREADS) gemm_group_kernel(GemmGroup g) {
  __shared__ double As[GT_K][GT_M + GT_LDS_PAD];
  __shared__ double Bs[GT_K][GT_N + GT_LDS_PAD];
  int tile = blockIdx.x;
  int pi = 0;
  while (pi + 1 < g.nprob && tile >= g.p[pi + 1].tile_begin) ++pi;
  const GemmProblem& P = g.p[pi];
  int local = tile - P.tile_begin;
  const int per_batch = P.tiles_m * P.tiles_n;
  const int bidx = local / per_batch;
  local -= bidx * per_batch;
  const int b1 = bidx % P.batch1, b2 = bidx / P.batch1;
  const int tm = local / P.tiles_n, tn = local % P.tiles_n;
  const int m0 = tm * GT_M, n0 = tn * GT_N;
  if (P.uplo == TRI_LOWER && n0 > m0 + GT_M - 1) return;
  if (P.uplo == TRI_UPPER && m0 > n0 + GT_N - 1) return;

  f64x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  for (int ti = 0; ti < P.nterms; ++ti) {
    const GemmTerm& t = P.t[ti];
    const double* A = t.A + b1 * t.sA1 + b2 * t.sA2;
    const double* B = t.B + b1 * t.sB1 + b2 * t.sB2;
    gemm_accumulate_term(t, A, B, P.M, P.N, m0, n0, As, Bs, acc);
  }

  double* C = P.C + b1 * P.sC1 + b2 * P.sC2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = m0 + wm + i * 16 + acc_row_f64(lane, r);
        int col = n0 + wn + j * 16 + acc_col(lane);
        if (row >= P.M || col >= P.N) continue;
        if (P.uplo == TRI_LOWER && col > row) continue;
        if (P.uplo == TRI_UPPER && col < row) continue;
        double* c = C + (int64_t)row * P.ldc + col;
        double v = acc[i][j][r];
        if (P.beta != 0.0) v += P.beta * (*c);
        *c = v;
      }
}

}  // namespace rocbta
