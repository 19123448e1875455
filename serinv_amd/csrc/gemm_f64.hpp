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
  // lower_packed (uplo == TRI_LOWER, tiles_m == tiles_n): a batch entry has
  // tiles_m (tiles_m + 1) / 2 workgroups, one per lower tile, instead of
  // tiles_m * tiles_n of which the upper ones would exit at once.
  int lower_packed;
  int tile_begin;  // first global tile index of this problem
  // Split-K: ksplit workgroups share one output tile, each over a contiguous
  // part of the K stages; partial tiles go to `partial` (ksplit * 4096
  // doubles per tile) and the last workgroup to finish a tile (per-tile
  // counter, zero on entry and reset on exit) sums them in split order.
  int ksplit;
  double* partial;
  int* counters;
  // Mixed precision: operands are read as FP64 and rounded to FP32 at the
  // MFMA, products accumulate in FP32 (v_mfma_f32_16x16x4_f32); the epilogue
  // (beta * C + result) runs in FP64.
  int f32;
  GemmTerm t[2];
};

struct GemmGroup {
  int nprob;
  int total_tiles;
  GemmProblem p[ROCBTA_MAX_PROBLEMS];
};

// ------------------------------------------------------------------ staging
// Two tile shapes: TS = 64 (64 x 64 output tile, K stage 32; four 16x16
// MFMA accumulators per wave) for small and many-problem launches, and
// TS = 128 (128 x 128 tile, K stage 16; sixteen accumulators per wave, one
// workgroup per CU) for large products, where the wider tile halves the
// operand traffic per flop and gives each wave 16 independent MFMA chains.
//
// A pipeline stage holds a TS x K slice of op(A) and of op(B)^T
// (rows m resp. n, columns k).  Each operand is read from global memory in
// its natural orientation with 16-byte loads (two doubles per lane, whole
// row segments per wave) and kept in LDS in that orientation:
//   K-contiguous source (A not transposed, B transposed): X[r][k], pitch PK
//   R-contiguous source (A transposed,     B not transposed): X[k][r], pitch PR
// An MFMA operand fetch (16 consecutive r at one k, four k per wave) then
// reads LDS with a small stride in either layout.
template <int TS>
struct GTile {
  static constexpr int K = (TS == 64) ? 32 : 16;
  static constexpr int PK = K + 2;
  static constexpr int PR = TS + 2;
  static constexpr int SLICE = (TS * PK > K * PR) ? TS * PK : K * PR;
  static constexpr int NB = TS / 32;  // 16x16 blocks per side of a wave's quadrant
};
constexpr int PK = GTile<64>::PK;
constexpr int PR = GTile<64>::PR;
constexpr int SLICE = GTile<64>::SLICE;
static_assert(GTile<64>::K == GT_K, "64-tile stage depth");

typedef double dbl2 __attribute__((ext_vector_type(2)));

struct StageRegs {
  dbl2 v[4];  // this thread's 8 elements of one operand slice
};

// Global -> registers.  ld: row stride of the stored matrix.  For a
// K-contiguous slice the stored rows are r (TS of them), columns k; for an
// R-contiguous slice the stored rows are k (K of them), columns r.
// rows_ok / cols_ok bound the stored row / column index (exclusive).
template <int TS, bool KCONT, bool FAST>
__device__ __forceinline__ void slice_load(const double* base, int64_t ld, int row0, int col0,
                                           int rows_ok, int cols_ok, StageRegs& r) {
  constexpr int TPR = KCONT ? GTile<TS>::K / 2 : TS / 2;  // threads per stored row
  const int tid = threadIdx.x & (GT_THREADS - 1);
  const int rsub = tid / TPR;
  const int c = (tid % TPR) * 2;
  constexpr int rstep = GT_THREADS / TPR;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = row0 + rsub + rstep * q, col = col0 + c;
    const double* src = base + (int64_t)row * ld + col;
    dbl2 x;
    if (FAST) {
      x = *reinterpret_cast<const dbl2*>(__builtin_assume_aligned(src, 16));
    } else {
      x.x = (row < rows_ok && col < cols_ok) ? src[0] : 0.0;
      x.y = (row < rows_ok && col + 1 < cols_ok) ? src[1] : 0.0;
    }
    r.v[q] = x;
  }
}

// Piece q (0..3) of this thread's part of a slice, to LDS.
template <int TS, bool KCONT>
__device__ __forceinline__ void slice_store_piece(const StageRegs& r, double* X, int q) {
  constexpr int TPR = KCONT ? GTile<TS>::K / 2 : TS / 2;
  const int tid = threadIdx.x & (GT_THREADS - 1);
  const int rsub = tid / TPR;
  const int c = (tid % TPR) * 2;
  constexpr int rstep = GT_THREADS / TPR;
  constexpr int pitch = KCONT ? GTile<TS>::PK : GTile<TS>::PR;
  *reinterpret_cast<dbl2*>(X + (rsub + rstep * q) * pitch + c) = r.v[q];
}

template <int TS, bool KCONT>
__device__ __forceinline__ void slice_store(const StageRegs& r, double* X) {
#pragma unroll
  for (int q = 0; q < 4; ++q) slice_store_piece<TS, KCONT>(r, X, q);
}

// Operand element (r, k) of a staged slice.
template <int TS, bool KCONT>
__device__ __forceinline__ double slice_at(const double* X, int r, int k) {
  return KCONT ? X[r * GTile<TS>::PK + k] : X[k * GTile<TS>::PR + r];
}

__device__ __forceinline__ void scale_acc(f64x4& a, double r) { a *= r; }
__device__ __forceinline__ void scale_acc(f32x4& a, double r) { a *= (float)r; }

// K range of a term restricted by the triangular shape of its operands.
template <int TS>
__device__ __forceinline__ void gemm_k_range(const GemmTerm& t, int m0, int n0, int& kbeg,
                                             int& kend) {
  kbeg = 0;
  kend = t.K;
  if (t.triA == TRI_LOWER) kend = min(kend, m0 + TS);
  if (t.triA == TRI_UPPER) kbeg = max(kbeg, m0);
  if (t.triB == TRI_LOWER) kbeg = max(kbeg, n0);
  if (t.triB == TRI_UPPER) kend = min(kend, n0 + TS);
  kbeg = (kbeg / GTile<TS>::K) * GTile<TS>::K;
}

__device__ __forceinline__ bool aligned16(const double* p, int64_t ld) {
  return ((reinterpret_cast<uintptr_t>(p) & 15) == 0) && ((ld & 1) == 0);
}

// One term's operands for one workgroup tile.  FAST: every stage is a full,
// 16-byte aligned TS x K slice (no bounds checks, vector loads only).
template <int TS, bool TA, bool TB, bool FAST>
struct TermView {
  const double* A;
  const double* B;
  int64_t lda, ldb;
  int M, N, m0, n0, ke;

  __device__ __forceinline__ void load(int k0, StageRegs& ra, StageRegs& rb) const {
    // A: K-contiguous when not transposed (stored M x K), else R-contiguous (stored K x M)
    if (!TA)
      slice_load<TS, true, FAST>(A, lda, m0, k0, M, ke, ra);
    else
      slice_load<TS, false, FAST>(A, lda, k0, m0, ke, M, ra);
    // B: op(B)(k, n); transposed => stored N x K (K-contiguous), else K x N
    if (TB)
      slice_load<TS, true, FAST>(B, ldb, n0, k0, N, ke, rb);
    else
      slice_load<TS, false, FAST>(B, ldb, k0, n0, ke, N, rb);
  }
  __device__ __forceinline__ void store(const StageRegs& ra, const StageRegs& rb, double* XA,
                                        double* XB) const {
    slice_store<TS, !TA>(ra, XA);
    slice_store<TS, TB>(rb, XB);
  }
};

// One LDS stage of MFMAs (wave quadrant (TS/2)^2 = NB x NB MFMA tiles).
// With ST, the next stage's slices (this thread's registers ra, rb) are
// written to the other LDS slot (SA, SB) during the MFMAs, one 16-byte piece
// after each group of a k step's MFMAs, instead of in a burst between
// stages: the LDS writes then overlap the matrix work.
template <int TS, bool TA, bool TB, typename ACC, bool ST = false>
__device__ __forceinline__ void gemm_mfma_stage(const double* XA, const double* XB,
                                                ACC (&acc)[GTile<TS>::NB][GTile<TS>::NB],
                                                const StageRegs* ra = nullptr,
                                                const StageRegs* rb = nullptr,
                                                double* SA = nullptr, double* SB = nullptr) {
  constexpr int NB = GTile<TS>::NB;
  const int lane = threadIdx.x & 63;
  const int wave = (threadIdx.x >> 6) & 3;
  const int wm = (wave >> 1) * (TS / 2), wn = (wave & 1) * (TS / 2);
  const int li = lane & 15, lk = lane >> 4;
  // operand registers double-buffered across k steps: the LDS reads of
  // step s + 1 are issued before the MFMAs of step s, which hide their
  // latency
  constexpr int NS = GTile<TS>::K / 4;
  double av[2][NB], bv[2][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) av[0][i] = slice_at<TS, !TA>(XA, wm + 16 * i + li, lk);
#pragma unroll
  for (int j = 0; j < NB; ++j) bv[0][j] = slice_at<TS, TB>(XB, wn + 16 * j + li, lk);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int cur = s & 1;
    if (s + 1 < NS) {
#pragma unroll
      for (int i = 0; i < NB; ++i)
        av[cur ^ 1][i] = slice_at<TS, !TA>(XA, wm + 16 * i + li, 4 * (s + 1) + lk);
#pragma unroll
      for (int j = 0; j < NB; ++j)
        bv[cur ^ 1][j] = slice_at<TS, TB>(XB, wn + 16 * j + li, 4 * (s + 1) + lk);
    }
    // keep the reads ahead of the MFMAs (the scheduler would otherwise sink
    // them to their use and wait on each)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) acc[i][j] = mfma_acc(av[cur][i], bv[cur][j], acc[i][j]);
    if (ST) {
      // pieces 0..3: A, 4..7: B, spread evenly over the NS k steps
#pragma unroll
      for (int q = s * 8 / NS; q < (s + 1) * 8 / NS; ++q) {
        if (q < 4)
          slice_store_piece<TS, !TA>(*ra, SA, q);
        else
          slice_store_piece<TS, TB>(*rb, SB, q - 4);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// acc += alpha * op(A) op(B) over the K range [kb, ke): LDS double buffer
// (slots 0/1) and two register stages in flight (x/y), unrolled by two so
// that every register set has a fixed role.  One barrier per stage: the
// MFMAs of stage it (one slot) write stage it+1's slices into the other slot
// as they go, and the barrier after them publishes those and frees the slot
// just read.  Measured on a 1024^3 product (64-tiles, 512 threads): 46.4 us
// against 55.9 us with the writes in a burst between two barriers.
//
// NH = 2 (512-thread workgroups): two 256-thread halves split the stages,
// half h taking stages h, h + 2, h + 4, ... with its own LDS double buffer.
// Both halves run the same number of iterations (a half without a stage in
// the last one only passes the barriers), so every __syncthreads is reached
// by all 512 threads.  Their accumulators are summed in gemm_tile_acc.
template <int TS, int NH, bool TA, bool TB, bool FAST, typename ACC>
__device__ __forceinline__ void gemm_term_pipeline(const TermView<TS, TA, TB, FAST>& v, int kb,
                                                   double* lds,
                                                   ACC (&acc)[GTile<TS>::NB][GTile<TS>::NB]) {
  constexpr int KS = GTile<TS>::K, SL = GTile<TS>::SLICE;
  const int ke = v.ke;
  const int nst = kb < ke ? (ke - kb + KS - 1) / KS : 0;
  if (nst == 0) return;
  const int h = (NH == 2) ? (int)(threadIdx.x >> 8) : 0;
  const int nit = (nst + NH - 1) / NH;  // iterations, the same in both halves
  auto has = [&](int it) { return it < nit && NH * it + h < nst; };
  auto kof = [&](int it) { return kb + (NH * it + h) * KS; };
  double* A0 = lds + h * 4 * SL;
  double* B0 = A0 + SL;
  double* A1 = A0 + 2 * SL;
  double* B1 = A0 + 3 * SL;
  StageRegs xa, xb, ya, yb;
  __syncthreads();  // LDS free (previous term / previous use)
  if (has(0)) v.load(kof(0), xa, xb);
  if (has(1)) v.load(kof(1), ya, yb);
  if (has(0)) v.store(xa, xb, A0, B0);
  __syncthreads();
  if (has(2)) v.load(kof(2), xa, xb);
  for (int it = 0; it < nit; it += 2) {
    // slot 0 holds iteration it; y holds it+1; x holds it+2 (in flight);
    // slot 1 is free (its last reader passed the previous barrier)
    if (has(it)) {
      if (has(it + 1))
        gemm_mfma_stage<TS, TA, TB, ACC, true>(A0, B0, acc, &ya, &yb, A1, B1);
      else
        gemm_mfma_stage<TS, TA, TB, ACC>(A0, B0, acc);
    }
    __syncthreads();
    if (has(it + 3)) v.load(kof(it + 3), ya, yb);
    if (it + 1 >= nit) break;
    if (has(it + 1)) {
      if (has(it + 2))
        gemm_mfma_stage<TS, TA, TB, ACC, true>(A1, B1, acc, &xa, &xb, A0, B0);
      else
        gemm_mfma_stage<TS, TA, TB, ACC>(A1, B1, acc);
    }
    __syncthreads();
    if (has(it + 4)) v.load(kof(it + 4), xa, xb);
  }
}

// Stages [sb, se) of the term (a stage = K consecutive k from the term's
// first k); the K range of those stages is clipped to the term's range.
template <int TS, int NH, bool TA, bool TB, typename ACC>
__device__ __forceinline__ void gemm_term(const GemmTerm& t, const double* A, const double* B,
                                          int M, int N, int m0, int n0, int sb, int se,
                                          double* lds, ACC (&acc)[GTile<TS>::NB][GTile<TS>::NB]) {
  constexpr int KS = GTile<TS>::K;
  int kb, ke;
  gemm_k_range<TS>(t, m0, n0, kb, ke);
  ke = min(ke, kb + se * KS);
  kb = kb + sb * KS;
  if (kb >= ke) return;
  const bool fast = aligned16(A, t.lda) && aligned16(B, t.ldb) && m0 + TS <= M &&
                    n0 + TS <= N && ((ke - kb) % KS) == 0;
  if (fast) {
    TermView<TS, TA, TB, true> v{A, B, t.lda, t.ldb, M, N, m0, n0, ke};
    gemm_term_pipeline<TS, NH, TA, TB, true, ACC>(v, kb, lds, acc);
  } else {
    TermView<TS, TA, TB, false> v{A, B, t.lda, t.ldb, M, N, m0, n0, ke};
    gemm_term_pipeline<TS, NH, TA, TB, false, ACC>(v, kb, lds, acc);
  }
}

// One TS x TS output tile (global tile index `tile`) of problem P, with
// accumulators of type ACC (f64x4: FP64, f32x4: mixed precision).  The
// problem's tiles_m / tiles_n / tile_begin were computed for this TS.
// Tile rows per group of the grouped tile order (gemm_tile_acc).  With
// xcd_tile's contiguous shares, a 1024^2 product (16 x 16 tiles) gives each
// XCD a 4 x 8 block: 4 A row panels and 8 B column panels in its L2 instead
// of 2 and all 16.  Measured: pobtasi at n=256, b=1024, a=32 42.4 -> 41.0 ms.
constexpr int GEMM_GROUP_M = 4;

template <int TS, int NH, typename ACC>
__device__ __forceinline__ void gemm_tile_acc(const GemmProblem& P, int tile, double* lds) {
  constexpr int NB = GTile<TS>::NB, KS = GTile<TS>::K;
  int local = (tile - P.tile_begin);
  const int split = local % P.ksplit;
  local /= P.ksplit;
  const int tile_id = local;  // tile index within the problem (for split-K bookkeeping)
  const int per_batch =
      P.lower_packed ? P.tiles_m * (P.tiles_m + 1) / 2 : P.tiles_m * P.tiles_n;
  const int bidx = local / per_batch;
  local -= bidx * per_batch;
  const int b1 = bidx % P.batch1, b2 = bidx / P.batch1;
  int tm, tn;
  if (P.lower_packed) {  // local = tm (tm + 1) / 2 + tn, 0 <= tn <= tm
    tm = (int)((__builtin_sqrtf(8.0f * (float)local + 1.0f) - 1.0f) * 0.5f);
    while ((tm + 1) * (tm + 2) / 2 <= local) ++tm;
    while (tm * (tm + 1) / 2 > local) --tm;
    tn = local - tm * (tm + 1) / 2;
  } else {
    // grouped order: GEMM_GROUP_M tile rows at a time, column-major within
    // the group, so that a run of consecutive tiles (one XCD's share, see
    // xcd_tile) is a block of rows x columns sharing both operand panels
    const int first_m = (local / (GEMM_GROUP_M * P.tiles_n)) * GEMM_GROUP_M;
    const int rows = min(GEMM_GROUP_M, P.tiles_m - first_m);
    const int within = local - first_m * P.tiles_n;
    tm = first_m + within % rows;
    tn = within / rows;
  }
  const int m0 = tm * TS, n0 = tn * TS;
  if (P.uplo == TRI_LOWER && n0 > m0 + TS - 1) return;
  if (P.uplo == TRI_UPPER && m0 > n0 + TS - 1) return;

  double* C = P.C + b1 * P.sC1 + b2 * P.sC2;
  const int lane = threadIdx.x & 63;
  const int wave = (threadIdx.x >> 6) & 3;
  const int wm = (wave >> 1) * (TS / 2), wn = (wave & 1) * (TS / 2);

  ACC acc[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = ACC{0, 0, 0, 0};

  // stage counts of the terms; this workgroup's share of the stage sequence
  int nst[2] = {0, 0};
  for (int ti = 0; ti < P.nterms; ++ti) {
    int kb, ke;
    gemm_k_range<TS>(P.t[ti], m0, n0, kb, ke);
    nst[ti] = kb < ke ? (ke - kb + KS - 1) / KS : 0;
  }
  // The terms' alphas are not applied to the operands (a product in the
  // staging path costs the wave a wait at the global load, measured 6% on a
  // 1024 x 1024 x 4096 product): the accumulators hold the sum in units of
  // a_ref, the last term's alpha (term 0's partial sum is scaled by
  // alpha_0 / a_ref when it is complete, exact for the +-1 in use), and the
  // epilogue multiplies by a_ref.  A second term with alpha 0 is skipped.
  const bool two = P.nterms == 2 && P.t[1].alpha != 0.0;
  const double a_ref = two ? P.t[1].alpha : P.t[0].alpha;
  const double r0 = two ? P.t[0].alpha / a_ref : 1.0;
  if (!two) nst[1] = 0;
  const int total = nst[0] + nst[1];
  const int g0 = (int)((int64_t)total * split / P.ksplit);
  const int g1 = (int)((int64_t)total * (split + 1) / P.ksplit);
  for (int ti = 0; ti < (two ? 2 : 1); ++ti) {
    const int off = ti ? nst[0] : 0;
    const int sb = max(g0 - off, 0), se = min(g1 - off, nst[ti]);
    if (ti == 1 && r0 != 1.0) {
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) scale_acc(acc[i][j], r0);
    }
    if (sb >= se) continue;
    const GemmTerm& t = P.t[ti];
    const double* A = t.A + b1 * t.sA1 + b2 * t.sA2;
    const double* B = t.B + b1 * t.sB1 + b2 * t.sB2;
    if (!t.transA && !t.transB)
      gemm_term<TS, NH, false, false, ACC>(t, A, B, P.M, P.N, m0, n0, sb, se, lds, acc);
    else if (!t.transA && t.transB)
      gemm_term<TS, NH, false, true, ACC>(t, A, B, P.M, P.N, m0, n0, sb, se, lds, acc);
    else if (t.transA && !t.transB)
      gemm_term<TS, NH, true, false, ACC>(t, A, B, P.M, P.N, m0, n0, sb, se, lds, acc);
    else
      gemm_term<TS, NH, true, true, ACC>(t, A, B, P.M, P.N, m0, n0, sb, se, lds, acc);
  }

  // NH = 2: the upper half adds its accumulators to the lower half's
  // through LDS; the lower half (threads 0..255) runs the epilogue, the upper
  // half only passes the remaining barriers.
  const bool lead = (NH == 1) || threadIdx.x < GT_THREADS;
  if (NH == 2) {
    __syncthreads();  // both pipelines are done with LDS
    const int tl = threadIdx.x & (GT_THREADS - 1);
    if (!lead) {
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) lds[((i * NB + j) * 4 + r) * GT_THREADS + tl] = acc[i][j][r];
    }
    __syncthreads();
    if (lead) {
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += lds[((i * NB + j) * 4 + r) * GT_THREADS + tl];
    }
  }

  if (P.ksplit > 1) {
    // Publish this split's partial tile; the last finisher sums all of them
    // in split order (deterministic) and runs the epilogue.
    const int tid = threadIdx.x & (GT_THREADS - 1);
    double* mine = P.partial + ((int64_t)tile_id * P.ksplit + split) * (TS * TS);
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (lead) mine[((i * NB + j) * 4 + r) * GT_THREADS + tid] = acc[i][j][r];
    // every wave releases its own partial stores before the barrier (the
    // barrier does not wait for them; thread 0's release covers its wave)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    __shared__ int s_last;
    if (threadIdx.x == 0) {
      const int old = __hip_atomic_fetch_add(&P.counters[tile_id], 1, __ATOMIC_ACQ_REL,
                                             __HIP_MEMORY_SCOPE_AGENT);
      s_last = (old == P.ksplit - 1);
      if (s_last) __hip_atomic_store(&P.counters[tile_id], 0, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (!s_last || !lead) return;
    const double* base = P.partial + (int64_t)tile_id * P.ksplit * (TS * TS);
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int idx = ((i * NB + j) * 4 + r) * GT_THREADS + tid;
          double sum = 0.0;
          for (int sp = 0; sp < P.ksplit; ++sp)
            sum += (sp == split) ? (double)acc[i][j][r] : base[(int64_t)sp * (TS * TS) + idx];
          acc[i][j][r] = sum;
        }
  }

  if (!lead) return;
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + acc_row(acc[i][j], lane, r);
        const int col = n0 + wn + j * 16 + acc_col(lane);
        if (row >= P.M || col >= P.N) continue;
        if (P.uplo == TRI_LOWER && col > row) continue;
        if (P.uplo == TRI_UPPER && col < row) continue;
        double* c = C + (int64_t)row * P.ldc + col;
        const double av = (double)acc[i][j][r];
        const double v = (P.beta != 0.0) ? a_ref * av + P.beta * (*c) : a_ref * av;
        *c = v;
        if (P.mirror && col != row) C[(int64_t)col * P.ldc + row] = v;
      }
}

template <int TS, int NH = 1>
__device__ __forceinline__ void gemm_tile(const GemmProblem& P, int tile, double* lds) {
  if (P.f32)
    gemm_tile_acc<TS, NH, f32x4>(P, tile, lds);
  else
    gemm_tile_acc<TS, NH, f64x4>(P, tile, lds);
}

// Workgroup -> tile order.  Workgroups are dealt to the 8 XCDs round-robin
// (workgroup w runs on XCD w mod 8), each with its own L2.  Giving XCD x a
// contiguous range of tile indices (row-major within a problem: whole tile
// rows) makes the workgroups of one XCD share their A row panels in that
// L2 instead of every XCD streaming every panel.  A bijection of
// [0, ntiles), so any dispatch order stays correct.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_tile(int w, int ntiles) {
  const int q = ntiles / N_XCD, r = ntiles % N_XCD;
  const int x = w % N_XCD, i = w / N_XCD;
  return x * q + min(x, r) + i;
}

__global__ void __launch_bounds__(GT_THREADS, 2) gemm_group_kernel(GemmGroup g) {
  __shared__ __attribute__((aligned(16))) double lds[4 * GTile<64>::SLICE];
  const int tile = xcd_tile(blockIdx.x, g.total_tiles);
  int pi = 0;
  while (pi + 1 < g.nprob && tile >= g.p[pi + 1].tile_begin) ++pi;
  gemm_tile<64>(g.p[pi], tile, lds);
}

// 64 x 64 tiles with 512 threads (one workgroup per CU, two waves per SIMD):
// the K stages of a tile are split between two 256-thread halves.  For
// launches with about one tile per CU this doubles the waves that issue
// MFMAs without a split-K pass through global memory.
__global__ void __launch_bounds__(2 * GT_THREADS, 1) gemm_group_kernel_h2(GemmGroup g) {
  __shared__ __attribute__((aligned(16))) double lds[8 * GTile<64>::SLICE];
  const int tile = xcd_tile(blockIdx.x, g.total_tiles);
  int pi = 0;
  while (pi + 1 < g.nprob && tile >= g.p[pi + 1].tile_begin) ++pi;
  gemm_tile<64, 2>(g.p[pi], tile, lds);
}

// 128 x 128 tiles for large products: one workgroup per CU, so that each
// wave can hold its sixteen accumulators without spilling.
__global__ void __launch_bounds__(GT_THREADS, 1) gemm_group_kernel_t128(GemmGroup g) {
  __shared__ __attribute__((aligned(16))) double lds[4 * GTile<128>::SLICE];
  const int tile = xcd_tile(blockIdx.x, g.total_tiles);
  int pi = 0;
  while (pi + 1 < g.nprob && tile >= g.p[pi + 1].tile_begin) ++pi;
  gemm_tile<128>(g.p[pi], tile, lds);
}

// The same over a device array of problems (sorted by tile_begin, 64-tiles):
// used by the batched plans, whose launch lists are built once and replayed.
__global__ void __launch_bounds__(GT_THREADS, 2) gemm_list_kernel(const GemmProblem* probs,
                                                                  int nprob) {
  __shared__ __attribute__((aligned(16))) double lds[4 * GTile<64>::SLICE];
  const int tile = blockIdx.x;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
  }
  gemm_tile<64>(probs[lo], tile, lds);
}

// The list form with 512-thread workgroups (the halves split each tile's K
// range, as in gemm_group_kernel_h2).
__global__ void __launch_bounds__(2 * GT_THREADS, 1) gemm_list_kernel_h2(const GemmProblem* probs,
                                                                         int nprob) {
  __shared__ __attribute__((aligned(16))) double lds[8 * GTile<64>::SLICE];
  const int tile = blockIdx.x;
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].tile_begin <= tile) lo = mid; else hi = mid - 1;
  }
  gemm_tile<64, 2>(probs[lo], tile, lds);
}

}  // namespace rocbta
