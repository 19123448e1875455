// Single-workgroup FP64 tile factorizations held in LDS (tiles up to 64x64).
//
//   potrf : L = chol(A) (lower, strict upper zero-filled), like
//           scipy.linalg.cholesky(lower=True) which the reference uses
//           (src/serinv/__init__.py:151).  A non-positive or NaN pivot at
//           tile column j records info = row_offset + j + 1 (LAPACK style).
//   trtri : X = L^{-1} for a lower-triangular L (strict upper zero-filled).
//
// Both are blocked by 16 columns.  The sequential part of a 16-column step
// (pivot, scaling, rank-1 updates inside the panel) runs in ONE wave with the
// panel rows spread over the 64 lanes and pivots broadcast by readlane, so it
// needs no workgroup barrier; the trailing updates of the tile are 16x16x4
// FP64 MFMA products spread over the four waves.
#pragma once
#include "common.hpp"

namespace rocbta {

constexpr int TT = 64;  // maximum tile order
constexpr int TT_THREADS = 256;
constexpr int TT_LD = TT + 1;
constexpr int PW = 16;  // panel width

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
  int mode;         // list launches: 0 potrf (+ inverse), 1 inverse only
  int grid_begin;   // list launches: first workgroup of this entry
};

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// d = sqrt(x), inv = 1/sqrt(x) for x > 0 (NaN propagates), from the hardware
// reciprocal square root y0 (relative error below 2^-22) and one third-order
// correction of it: with e = 1 - x y0^2,
//   1/sqrt(x) = y0 (1 - e)^(-1/2) = y0 (1 + e/2 + 3e^2/8 + O(e^3)),
// so y = y0 + (y0 e)(1/2 + 3e/8) leaves a relative error of about 2.5 * 2^-66
// before rounding (four dependent operations after the seed; two Newton
// steps take six).  The square root is one Newton step on d = x y,
// d <- d + y (x - d^2) / 2, which is not on the path of inv.
__device__ __forceinline__ void pivot_sqrt(double x, double& d, double& inv) {
  const double y0 = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-(x * y0), y0, 1.0);
  const double p = __builtin_fma(0.375, e, 0.5);
  const double y = __builtin_fma(y0 * e, p, y0);
  double s = x * y;
  const double rr = __builtin_fma(-s, s, x);
  s = __builtin_fma(0.5 * y, rr, s);
  d = s;
  inv = y;
}

// Unblocked Cholesky of the panel S[p:m, p:p+w] (w <= 16, m - p <= 64) by
// the calling wave; lane l owns row p + l.  Lower entries are written back.
template <int WFIX>
__device__ __forceinline__ void wave_panel_potrf(double (*S)[TT_LD], int p, int m, int w_rt,
                                                 int* info, int64_t row0) {
  const int w = WFIX ? WFIX : w_rt;
  const int lane = threadIdx.x & 63;
  const int i = p + lane;
  const bool act = i < m;
  double r[PW];
#pragma unroll
  for (int k = 0; k < PW; ++k) r[k] = (act && k < w) ? S[i][p + k] : 0.0;
#pragma unroll
  for (int j = 0; j < PW; ++j) {
    if (j < w) {
      double piv = readlane_f64(r[j], j);
      if (!(piv > 0.0)) {
        if (lane == 0) record_info(info, (int)(row0 + p + j + 1));
        piv = __builtin_nan("");
      }
      double d, inv;
      pivot_sqrt(piv, d, inv);
      // the column is broadcast from its scaled value; the diagonal entry d
      // (lane j) is not on the path to the next pivot
      const double sc = r[j] * inv;
#pragma unroll
      for (int k = j + 1; k < PW; ++k) {
        if (k < w) {
          const double lkj = readlane_f64(sc, k);
          r[k] -= sc * lkj;
        }
      }
      r[j] = (lane == j) ? d : (lane > j ? sc : r[j]);
    }
  }
#pragma unroll
  for (int k = 0; k < PW; ++k)
    if (act && k < w && p + k <= i) S[i][p + k] = r[k];
}

// Full-width (16-column) form of the panel Cholesky, p + 16 <= m.  The wave
// is bound by instruction issue, and most of the issue went into readlane
// broadcasts of the row entries L(k,j) (two per double).  Here the panel is
// cut into four 4-column sub-panels:
//   * inside a sub-panel, right-looking with readlane broadcasts (at most
//     3 values per pivot);
//   * a finished sub-panel is written to S; the rows k of later sub-panels
//     read their four entries L(k, c0..c0+3) back as LDS broadcasts (one
//     address for all lanes), except the next pivot column c0+4, which
//     takes readlanes because it is on the dependency path.
// A non-positive or NaN pivot turns into NaN through the reciprocal square
// root, so the loop has no branch; the first failing column is recorded
// once after it.  Lanes above the diagonal of a column hold values that
// are never read or written back.
__device__ __forceinline__ void wave_panel_potrf16(double (*S)[TT_LD], int p, int m, int* info,
                                                   int64_t row0) {
  const int lane = threadIdx.x & 63;
  const int i = p + lane;
  const bool act = i < m;
  const int ic = act ? i : p;  // clamped row for the loads
  double r[PW];
#pragma unroll
  for (int k = 0; k < PW; ++k) r[k] = act ? S[ic][p + k] : 0.0;
  int bad = PW;
#pragma unroll
  for (int c0 = 0; c0 < PW; c0 += 4) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int j = c0 + jj;
      const double piv = readlane_f64(r[j], j);
      bad = (bad == PW && !(piv > 0.0)) ? j : bad;
      double d, inv;
      pivot_sqrt(piv, d, inv);
      // broadcasts read the scaled column before the diagonal entry d is
      // selected in (d is off the path to the next pivot)
      const double sc = r[j] * inv;
#pragma unroll
      for (int k = j + 1; k < c0 + 4; ++k) r[k] -= sc * readlane_f64(sc, k);
      r[j] = (lane == j) ? d : sc;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (act && p + c0 + q <= i) S[i][p + c0 + q] = r[c0 + q];
    if (c0 + 4 < PW) {
      __builtin_amdgcn_wave_barrier();  // the sub-panel's LDS stores precede the reads below
      // Every broadcast read of the finished rows is issued before the next
      // pivot column's update, which takes readlanes: the reads' latency
      // then overlaps it instead of stalling the pivot path read by read.
      double lrow[PW - 5][4];
#pragma unroll
      for (int k = c0 + 5; k < PW; ++k)
#pragma unroll
        for (int q = 0; q < 4; ++q) lrow[k - c0 - 5][q] = S[p + k][p + c0 + q];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 4; ++q) r[c0 + 4] -= r[c0 + q] * readlane_f64(r[c0 + q], c0 + 4);
#pragma unroll
      for (int k = c0 + 5; k < PW; ++k)
#pragma unroll
        for (int q = 0; q < 4; ++q) r[k] -= r[c0 + q] * lrow[k - c0 - 5][q];
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (bad < PW && lane == 0) record_info(info, (int)(row0 + p + bad + 1));
}

// S[q:m, q:m] (lower part) -= P P^T with P = S[q:m, p:p+w], on MFMA, all
// waves.  Output 16x16 blocks are dealt round-robin to the waves.
template <int WFIX>
__device__ __forceinline__ void lds_syrk_update(double (*S)[TT_LD], int p, int w_rt, int q,
                                                int m) {
  const int w = WFIX ? WFIX : w_rt;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nb = (m - q + 15) / 16;
  const int li = lane & 15, lk = lane >> 4;
  int t = 0;
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = 0; bj <= bi; ++bj, ++t) {
      if ((t & 3) != wave) continue;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
      const int ra = q + bi * 16 + li, rb = q + bj * 16 + li;
#pragma unroll
      for (int kk = 0; kk < PW; kk += 4) {
        const int k = kk + lk;
        const double av = (k < w && ra < m) ? S[ra][p + k] : 0.0;
        const double bv = (k < w && rb < m) ? S[rb][p + k] : 0.0;
        acc = mfma_f64(av, bv, acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = q + bi * 16 + acc_row_f64(lane, r);
        const int col = q + bj * 16 + acc_col(lane);
        if (row < m && col <= row) S[row][col] -= acc[r];
      }
    }
}

// Full 64-tile form: S[c0:64, c0:c1] (lower part) -= P P^T restricted to
// the 16-wide column blocks [c0, c1), P = S[:, p:p+16].  Block t of the
// enumeration goes to wave wave0 + t % nw; other waves return at once.
__device__ __forceinline__ void lds_syrk_cols(double (*S)[TT_LD], int p, int c0, int c1,
                                              int wave0, int nw) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  int t = 0;
  for (int cb = c0 / 16; cb < c1 / 16; ++cb)
    for (int rb = cb; rb < TT / 16; ++rb, ++t) {
      if (wave != wave0 + t % nw) continue;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
      const int ra = rb * 16 + li, rc = cb * 16 + li;
#pragma unroll
      for (int kk = 0; kk < PW; kk += 4) acc = mfma_f64(S[ra][p + kk + lk], S[rc][p + kk + lk], acc);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + acc_row_f64(lane, r), col = cb * 16 + acc_col(lane);
        if (col <= row) S[row][col] -= acc[r];
      }
    }
}

struct NoIdle {
  __device__ void operator()(int) const {}
};

// Blocked Cholesky of S (m x m, lower part valid) in place.  MFIX = 64
// compiles the full-tile case with all sizes constant and a look-ahead: after
// panel p only the next panel's columns are updated before the next panel
// is factored (wave 0), while waves 1-3 apply panel p to the columns beyond
// it.  Full tiles: after that update (if any), waves 1-3 run idle(p) while
// wave 0 factors panel p + 16; panels 0 .. p are final by then.  The call
// ends with a barrier after the last of them.  MFIX = 0 handles any m <= 64.
template <int MFIX, typename IDLE = NoIdle>
__device__ __forceinline__ void lds_potrf_t(double (*S)[TT_LD], int m_rt, int* info,
                                            int64_t row0, IDLE idle = IDLE()) {
  const int m = MFIX ? MFIX : m_rt;
  const int wave = threadIdx.x >> 6;
  if (MFIX == TT) {
    if (wave == 0) wave_panel_potrf16(S, 0, TT, info, row0);
    __syncthreads();
#pragma unroll
    for (int p = 0; p + PW < TT; p += PW) {
      lds_syrk_cols(S, p, p + PW, p + 2 * PW, 0, 4);  // next panel's columns
      __syncthreads();
      if (wave == 0)
        wave_panel_potrf16(S, p + PW, TT, info, row0);
      else {
        if (p + 2 * PW < TT) lds_syrk_cols(S, p, p + 2 * PW, TT, 1, 3);  // the rest, beside the panel
        idle(p);
      }
      __syncthreads();
    }
    return;
  }
#pragma unroll
  for (int p = 0; p < m; p += PW) {
    const int w = min(PW, m - p);
    if (MFIX || w == PW) {
      if (wave == 0) wave_panel_potrf16(S, p, m, info, row0);
    } else {
      if (wave == 0) wave_panel_potrf<0>(S, p, m, w, info, row0);
    }
    __syncthreads();
    if (p + w < m) {
      if (MFIX)
        lds_syrk_update<PW>(S, p, PW, p + w, m);
      else
        lds_syrk_update<0>(S, p, w, p + w, m);
      __syncthreads();
    }
  }
}

__device__ __forceinline__ void lds_potrf(double (*S)[TT_LD], int m, int* info, int64_t row0) {
  if (m == TT)
    lds_potrf_t<TT>(S, TT, info, row0);
  else
    lds_potrf_t<0>(S, m, info, row0);
}

// X = L^{-1} (m x m lower, in S) into Xs (lower; strict upper zero).
// Step 1: each wave inverts 16x16 diagonal blocks by Gauss-Jordan sweeps,
// lane = (row i, column group g of four columns).  Step 2: off-diagonal
// blocks by increasing distance d from the diagonal on MFMA:
//   X_IJ = -X_II * sum_{K=J}^{I-1} L_IK X_KJ.
// Ws: 64 x 17 scratch; at distance d the block (I, I-d) of W sits in rows
// 16I .. 16I+15.
constexpr int WS_LD = 17;

// Inverse of the full 16x16 lower-triangular block of S at (o, o) by the
// calling wave, written to Xs as a full block (zeros above the diagonal).
// Lane c (and its copies c + 16, c + 32, c + 48) forms column c by forward
// substitution, x_i = (delta_ic - sum_{k<i} L_ik x_k) / L_ii; the entries
// L_ik are LDS broadcasts and 1 / L_ii comes from lane i by readlane, so the
// step needs no division on its dependency path and no cross-lane shuffle.
// The block is written to dst (row pitch ld).
__device__ __forceinline__ void wave_diag_inv16_to(const double (*S)[TT_LD], int o, double* dst,
                                                   int ld) {
  const int lane = threadIdx.x & 63;
  const int c = lane & 15;
  const double dinv = 1.0 / S[o + c][o + c];
  const double* S0 = &S[0][0];
  double x[16];
  // The LDS broadcast reads of row i+1, L(i+1, 0..i), are issued while row i
  // is substituted, so each row's reads have landed when it starts.  Left to
  // itself the compiler either waits after every read or hoists every row's
  // reads to the top (all 120 entries live at once).  The read address of
  // row i+1 takes a zero produced from x[i-1] (an asm move with that input),
  // which keeps the reads below x[i-1]; the scheduling barrier keeps them
  // above row i's work.
  double cur[16], nxt[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    if (i + 1 < 16) {
      int zero = 0;
      if (i >= 1) asm volatile("v_mov_b32 %0, 0" : "=v"(zero) : "v"(x[i - 1]));
      const double* row = S0 + (o + i + 1) * TT_LD + o + zero;
#pragma unroll
      for (int k = 0; k <= i; ++k) nxt[k] = row[k];
    }
    __builtin_amdgcn_sched_barrier(0);
    double acc = (i == c) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) acc -= cur[k] * x[k];
    x[i] = acc * readlane_f64(dinv, i);
#pragma unroll
    for (int k = 0; k <= i; ++k) cur[k] = nxt[k];
  }
  if (lane < 16) {
#pragma unroll
    for (int i = 0; i < 16; ++i) dst[i * ld + c] = x[i];
  }
}
__device__ __forceinline__ void wave_diag_inv16(const double (*S)[TT_LD], double (*Xs)[TT_LD],
                                                int o) {
  wave_diag_inv16_to(S, o, &Xs[o][o], TT_LD);
}

// Off-diagonal 16x16 block of X = L^{-1} by the calling wave (J < I):
//   X_IJ = -X_II * W,  W = sum_{K=J}^{I-1} L_IK X_KJ,
// with L in S, the blocks X_KJ (K < I) in Xs, and X_II read from XI (row
// pitch ldi).  X_IJ is written to Xs.  Lane (q, j) = (lane >> 4, lane & 15)
// holds W(4r + q, j) in w[r] after the first product, which is the B operand
// of k-step r of the second, so W never leaves the registers.
__device__ __forceinline__ void wave_trtri_offdiag(const double (*S)[TT_LD], double (*Xs)[TT_LD],
                                                   const double* XI, int ldi, int I, int J) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  f64x4 w = {0.0, 0.0, 0.0, 0.0};
  for (int K = J; K < I; ++K)
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      const int k = K * 16 + kk + lk;
      w = mfma_f64(S[I * 16 + li][k], Xs[k][J * 16 + li], w);
    }
  f64x4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < 4; ++r) x = mfma_f64(XI[li * ldi + 4 * r + lk], w[r], x);
#pragma unroll
  for (int r = 0; r < 4; ++r) Xs[I * 16 + acc_row_f64(lane, r)][J * 16 + acc_col(lane)] = -x[r];
}

// ZERO = false: the caller guarantees that the blocks of Xs strictly above
// the 16x16 block diagonal already hold zeros (they are never written here).
template <int MFIX, bool ZERO = true>
__device__ __forceinline__ void lds_trtri_t(const double (*S)[TT_LD], double (*Xs)[TT_LD],
                                            double (*Ws)[WS_LD], int m_rt) {
  const int m = MFIX ? MFIX : m_rt;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tid = threadIdx.x;
  if (ZERO || !MFIX) {
    for (int idx = tid; idx < TT * TT; idx += TT_THREADS) Xs[idx / TT][idx % TT] = 0.0;
    __syncthreads();
  }
  const int nb = (m + 15) / 16;
  if (MFIX == TT) {
    wave_diag_inv16(S, Xs, __builtin_amdgcn_readfirstlane(16 * wave));
  } else
  for (int blk = wave; blk < nb; blk += 4) {
    const int o = blk * 16, w = min(16, m - o);
    const int i = lane & 15, g = lane >> 4;
    double x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = (i == g * 4 + q) ? 1.0 : 0.0;
    for (int j = 0; j < w; ++j) {
      if (i == j) {
        const double inv = 1.0 / S[o + j][o + j];
#pragma unroll
        for (int q = 0; q < 4; ++q) x[q] *= inv;
      }
      const double lij = (i > j && i < w) ? S[o + i][o + j] : 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double xj = __shfl(x[q], g * 16 + j, 64);
        x[q] -= lij * xj;
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = g * 4 + q;
      if (i < w && c < w && c <= i) Xs[o + i][o + c] = x[q];
    }
  }
  __syncthreads();
  const int li = lane & 15, lk = lane >> 4;
  for (int d = 1; d < nb; ++d) {
    // W_J = sum_{K=J}^{J+d-1} L_{J+d,K} X_{K,J}  (into Ws block row J+d, col J)
    for (int J = wave; J + d < nb; J += 4) {
      const int I = J + d;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
      for (int K = J; K < I; ++K)
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
          const int k = K * 16 + kk + lk;
          const int ra = I * 16 + li, cb = J * 16 + li;
          const double av = (ra < m && k < m) ? S[ra][k] : 0.0;
          const double bv = (k < m && cb < m) ? Xs[k][cb] : 0.0;
          acc = mfma_f64(av, bv, acc);
        }
#pragma unroll
      for (int r = 0; r < 4; ++r) Ws[I * 16 + acc_row_f64(lane, r)][acc_col(lane)] = acc[r];
    }
    __syncthreads();
    for (int J = wave; J + d < nb; J += 4) {
      const int I = J + d;
      f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 16; kk += 4) {
        const int k = I * 16 + kk + lk;
        const int ra = I * 16 + li;
        const double av = (ra < m && k < m) ? Xs[ra][k] : 0.0;
        const double bv = (k < m) ? Ws[k][li] : 0.0;
        acc = mfma_f64(av, bv, acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = I * 16 + acc_row_f64(lane, r), col = J * 16 + acc_col(lane);
        if (row < m && col < m) Xs[row][col] = -acc[r];
      }
    }
    __syncthreads();
  }
}

template <bool ZERO = true>
__device__ __forceinline__ void lds_trtri(const double (*S)[TT_LD], double (*Xs)[TT_LD],
                                          double (*Ws)[WS_LD], int m) {
  if (m == TT)
    lds_trtri_t<TT, ZERO>(S, Xs, Ws, TT);
  else
    lds_trtri_t<0>(S, Xs, Ws, m);
}

// mode 0: potrf (+ inverse into X if X != nullptr); mode 1: trtri only.
template <int MODE>
__device__ __forceinline__ void tile_factor_body(const TileBatch& tb, int bz, double (*S)[TT_LD],
                                                 double (*Xs)[TT_LD], double (*Ws)[WS_LD]) {
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
    lds_trtri(S, Xs, Ws, m);
    double* X = tb.X + b1 * tb.sX1 + b2 * tb.sX2;
    for (int idx = tid; idx < m * m; idx += TT_THREADS) {
      int r = idx / m, c = idx % m;
      X[(int64_t)r * tb.ldx + c] = Xs[r][c];
    }
  }
}

template <int MODE>
__global__ void __launch_bounds__(TT_THREADS) tile_factor_kernel(TileBatch tb) {
  __shared__ double S[TT][TT_LD];
  __shared__ double Xs[TT][TT_LD];
  __shared__ double Ws[TT][WS_LD];
  tile_factor_body<MODE>(tb, blockIdx.x, S, Xs, Ws);
}

// The same over a device array of entries sorted by grid_begin (batched
// plans); each entry runs in its own mode.
__global__ void __launch_bounds__(TT_THREADS) tile_factor_list_kernel(const TileBatch* tbs,
                                                                      int ntb) {
  __shared__ double S[TT][TT_LD];
  __shared__ double Xs[TT][TT_LD];
  __shared__ double Ws[TT][WS_LD];
  const int wg = blockIdx.x;
  int lo = 0, hi = ntb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tbs[mid].grid_begin <= wg) lo = mid; else hi = mid - 1;
  }
  const TileBatch& tb = tbs[lo];
  if (tb.mode == 0)
    tile_factor_body<0>(tb, wg - tb.grid_begin, S, Xs, Ws);
  else
    tile_factor_body<1>(tb, wg - tb.grid_begin, S, Xs, Ws);
}

}  // namespace rocbta
