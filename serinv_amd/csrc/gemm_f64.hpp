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
constexpr int GT_K = 32;   // K slice staged through LDS per pipeline stage
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
  int uplo;    // TRI_NONE: full C; TRI_LOWER: only row >= col is written
  int mirror;  // with uplo == TRI_LOWER: also write C[col][row] = C[row][col]
  int tiles_m, tiles_n;
  int tile_begin;  // first global tile index of this problem
  GemmTerm t[2];
};

struct GemmGroup {
  int nprob;
  int total_tiles;
  GemmProblem p[ROCBTA_MAX_PROBLEMS];
};

// One pipeline stage: alpha*op(A) (64 x GT_K) and op(B) (GT_K x 64) of one
// term, 8 + 8 doubles per thread.  Out-of-range elements are zero.
struct StageRegs {
  double a[8];
  double b[8];
};

// Register <-> LDS mapping of a stage (k-major LDS: As[k][m], Bs[k][n]):
//   A not transposed (row-major M x K): thread t: m = t >> 2, k = 8*(t&3) .. +7
//   A transposed (stored K x M):        thread t: k = t >> 3, m = 8*(t&7) .. +7
//   B not transposed (stored K x N):    thread t: k = t >> 3, n = 8*(t&7) .. +7
//   B transposed (stored N x K):        thread t: n = t >> 2, k = 8*(t&3) .. +7
__device__ __forceinline__ void gemm_load_stage(const GemmTerm& t, const double* A,
                                                const double* B, int M, int N, int m0, int n0,
                                                int k0, int kend, StageRegs& r) {
  const int tid = threadIdx.x;
  if (!t.transA) {
    const int m = m0 + (tid >> 2), kq = k0 + (tid & 3) * 8;
    const double* src = A + (int64_t)m * t.lda + kq;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.a[j] = (m < M && kq + j < kend) ? t.alpha * src[j] : 0.0;
  } else {
    const int k = k0 + (tid >> 3), mq = m0 + (tid & 7) * 8;
    const double* src = A + (int64_t)k * t.lda + mq;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.a[j] = (k < kend && mq + j < M) ? t.alpha * src[j] : 0.0;
  }
  if (!t.transB) {
    const int k = k0 + (tid >> 3), nq = n0 + (tid & 7) * 8;
    const double* src = B + (int64_t)k * t.ldb + nq;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.b[j] = (k < kend && nq + j < N) ? src[j] : 0.0;
  } else {
    const int n = n0 + (tid >> 2), kq = k0 + (tid & 3) * 8;
    const double* src = B + (int64_t)n * t.ldb + kq;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.b[j] = (n < N && kq + j < kend) ? src[j] : 0.0;
  }
}

typedef double GemmLds[GT_K][GT_M + GT_LDS_PAD];

__device__ __forceinline__ void gemm_store_stage(const GemmTerm& t, const StageRegs& r,
                                                 GemmLds& As, GemmLds& Bs) {
  const int tid = threadIdx.x;
  if (!t.transA) {
    const int m = tid >> 2, kq = (tid & 3) * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) As[kq + j][m] = r.a[j];
  } else {
    const int k = tid >> 3, mq = (tid & 7) * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) As[k][mq + j] = r.a[j];
  }
  if (!t.transB) {
    const int k = tid >> 3, nq = (tid & 7) * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) Bs[k][nq + j] = r.b[j];
  } else {
    const int n = tid >> 2, kq = (tid & 3) * 8;
#pragma unroll
    for (int j = 0; j < 8; ++j) Bs[kq + j][n] = r.b[j];
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

__device__ __forceinline__ void gemm_mfma_stage(const GemmLds& As, const GemmLds& Bs,
                                                f64x4 (&acc)[2][2]) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < GT_K; kk += 4) {
    const double a0 = As[kk + lk][wm + li];
    const double a1 = As[kk + lk][wm + 16 + li];
    const double b0 = Bs[kk + lk][wn + li];
    const double b1 = Bs[kk + lk][wn + 16 + li];
    acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
    acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
    acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
    acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
  }
}

// Flat stage sequence over the (at most two) terms of a problem.
struct StageSeq {
  int nst[2];
  int kb[2], ke[2];
};

__device__ __forceinline__ void stage_locate(const StageSeq& q, int s, int& term, int& k0) {
  term = (s < q.nst[0]) ? 0 : 1;
  const int ls = term ? s - q.nst[0] : s;
  k0 = q.kb[term] + ls * GT_K;
}

__global__ void __launch_bounds__(GT_THREADS) gemm_group_kernel(GemmGroup g) {
  __shared__ GemmLds As[2];
  __shared__ GemmLds Bs[2];
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

  const double* Ab[2];
  const double* Bb[2];
  StageSeq q{};
  for (int ti = 0; ti < 2; ++ti) {
    q.nst[ti] = 0;
    if (ti < P.nterms) {
      const GemmTerm& t = P.t[ti];
      Ab[ti] = t.A + b1 * t.sA1 + b2 * t.sA2;
      Bb[ti] = t.B + b1 * t.sB1 + b2 * t.sB2;
      gemm_k_range(t, m0, n0, q.kb[ti], q.ke[ti]);
      q.nst[ti] = q.kb[ti] < q.ke[ti] ? (q.ke[ti] - q.kb[ti] + GT_K - 1) / GT_K : 0;
    }
  }
  const int nst = q.nst[0] + q.nst[1];

  double* C = P.C + b1 * P.sC1 + b2 * P.sC2;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;

  // Prefetch the C tile early (its latency then overlaps the main loop).
  f64x4 cin[2][2];
  const bool readc = P.beta != 0.0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + acc_row_f64(lane, r);
        const int col = n0 + wn + j * 16 + acc_col(lane);
        cin[i][j][r] = (readc && row < P.M && col < P.N) ? C[(int64_t)row * P.ldc + col] : 0.0;
      }

  f64x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  // Pipeline: LDS double buffer + two register stages in flight.
  StageRegs r0, r1;
  int term, k0;
  if (nst > 0) {
    stage_locate(q, 0, term, k0);
    gemm_load_stage(P.t[term], Ab[term], Bb[term], P.M, P.N, m0, n0, k0, q.ke[term], r0);
    gemm_store_stage(P.t[term], r0, As[0], Bs[0]);
  }
  if (nst > 1) {
    stage_locate(q, 1, term, k0);
    gemm_load_stage(P.t[term], Ab[term], Bb[term], P.M, P.N, m0, n0, k0, q.ke[term], r1);
  }
  if (nst > 2) {
    stage_locate(q, 2, term, k0);
    gemm_load_stage(P.t[term], Ab[term], Bb[term], P.M, P.N, m0, n0, k0, q.ke[term], r0);
  }
  __syncthreads();
  for (int s = 0; s < nst; ++s) {
    gemm_mfma_stage(As[s & 1], Bs[s & 1], acc);
    if (s + 1 < nst) {
      stage_locate(q, s + 1, term, k0);
      gemm_store_stage(P.t[term], (s & 1) ? r0 : r1, As[(s + 1) & 1], Bs[(s + 1) & 1]);
    }
    if (s + 3 < nst) {
      stage_locate(q, s + 3, term, k0);
      gemm_load_stage(P.t[term], Ab[term], Bb[term], P.M, P.N, m0, n0, k0, q.ke[term],
                      (s & 1) ? r0 : r1);
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + acc_row_f64(lane, r);
        const int col = n0 + wn + j * 16 + acc_col(lane);
        if (row >= P.M || col >= P.N) continue;
        if (P.uplo == TRI_LOWER && col > row) continue;
        if (P.uplo == TRI_UPPER && col < row) continue;
        const double v = acc[i][j][r] + P.beta * cin[i][j][r];
        C[(int64_t)row * P.ldc + col] = v;
        if (P.mirror && col != row) C[(int64_t)col * P.ldc + row] = v;
      }
}

}  // namespace rocbta
