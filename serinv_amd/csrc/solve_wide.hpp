// Wide block sweep of pobtas for many right-hand sides (gfx950, FP64 MFMA).
//
// Reference recurrence: src/serinv/algs/pobtas.py:80-126 (forward) and
// :127-159 (backward), restricted to the block-tridiagonal part; the arrow
// and tip terms run as batched products around the launch.  With the
// inverses Linv_i of the diagonal blocks (as in solve_block.hpp):
//
//   forward,  i = 0 .. n-1:   t_i = b_i - L_{i,i-1} y_{i-1}      (phase A)
//                             y_i = Linv_i t_i                    (phase B)
//   backward, i = n-1 .. 0:   s_i = y'_i - L_{i+1,i}^T x_{i+1}    (phase A)
//                             x_i = Linv_i^T s_i                  (phase B)
//
// (y'_i = y_i - R_i^T x_t is formed before the backward launch.)  One
// launch takes up to WD_CW = 64 right-hand-side columns, so the factor is
// read once per 64 columns instead of once per 8 as in the block sweep.
// The backward products read L_{i+1,i} and Linv_i by columns, so no
// transposed copies are needed.
//
// Work split: an item is a 16 x 16 output tile (16 rows of a block, 16 of
// the columns); workgroup w takes items w, w + G, ... of every phase.  The
// WD_NW waves of a workgroup split the reduction index in groups of 16 and
// accumulate on v_mfma_f64_16x16x4f64; their partial tiles are added
// through LDS in wave order, so results do not depend on scheduling.
// Within a group of 16 reduction indices, lane l (row or column l % 16,
// quarter g = l / 16) takes indices 4g .. 4g + 3 on the group's four MFMA
// steps: one 32-byte read of the factor row per lane per group.  A wave
// takes every WD_NW-th group and issues the loads of WD_NB of them before it
// polls any, so that a phase costs one or two memory round trips, not one
// per group.
//
// Published rows (t, y / s, x) live in buffers pre-filled with SW_EMPTY and
// are stored with relaxed agent-scope atomic stores; consumers poll the
// values themselves (solve_sweep.hpp): the data is its own flag, so a read
// costs one round trip.  (Plain stores with a per-tile flag, release and
// acquire, as in pobtaf_chain.hpp, were measured slower here: r = 64 at
// n=256, b=1024, N+C 31.5 ms against 25.5 ms.)  Every wait is bounded; an
// overrun sets *status = -1 and later waits return at once.  Items are
// taken in phase order (A of block i, B of block i, A of the next block),
// and every wait targets a value of an earlier phase, so a grid of
// co-resident workgroups always progresses.
#pragma once
#include "solve_block.hpp"

namespace rocbta {

constexpr int WD_CW = 64;       // right-hand-side columns per launch
constexpr int WD_T = 16;        // output tile: WD_T rows x WD_T columns
constexpr int WD_NW = 8;        // waves per workgroup (they split the reduction)
constexpr int WD_THREADS = 64 * WD_NW;
constexpr int WD_MAX_G = 1024;
constexpr int WD_NB = 8;        // K groups of 16 whose loads one wave issues together

struct WideSweepP {
  const double* L;     // L_{i+1,i} at i b^2 (backward reads it by columns)
  const double* Linv;  // Linv_i (zero above; backward reads it by columns)
  double* B;           // right-hand sides, columns [0, r), row stride ldb
  int64_t ldb;
  int n, b, r;         // b % WD_T == 0, r <= WD_CW
  int trans;
  int G;
  double* sbuf;        // (n*b, WD_CW): published solution rows
  double* tbuf;        // (n*b, WD_CW): published phase-A rows
  int* status;
  long long spin_limit;
};

__device__ __forceinline__ bool wd_spin(const WideSweepP& P, long long& spins) {
  if (++spins > P.spin_limit) {
    __hip_atomic_store(P.status, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
  }
  if ((spins & 63) == 0 &&
      __hip_atomic_load(P.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0)
    return false;
  __builtin_amdgcn_s_sleep(1);
  return true;
}

// sum_j op(M)[k0 + row][j] * src[j][c0 + col] over j in [jlo, jhi) (a
// multiple of 16) for the 16 x 16 tile, into red (the waves' partials) and
// then out[row * 16 + col] (wave-ordered sum).  M has row stride b; op(M) =
// M, or M^T read by columns (TR: for a reduction index, the 16 lanes of a
// quarter-wave read 128 contiguous bytes of row j); src rows are polled
// (row stride WD_CW).  Ends with a barrier.
template <bool TR>
__device__ __forceinline__ void wd_tile(const WideSweepP& P, const double* M, int k0,
                                        const double* src, int c0, int jlo, int jhi, double* red,
                                        double* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, g4 = (lane >> 4) * 4;
  const double* mrow = M + (int64_t)(k0 + li) * P.b;  // TR: M^T's row k0 + li is column k0 + li
  f64x4 acc = {0.0, 0.0, 0.0, 0.0};
  // WD_NB groups per round: every load of the round is issued before the
  // first poll, so a round costs one memory round trip
  for (int jr = jlo + 16 * wave; jr < jhi; jr += 16 * WD_NW * WD_NB) {
    double m[WD_NB][4];
    unsigned long long v[WD_NB][4];
#pragma unroll
    for (int u = 0; u < WD_NB; ++u) {
      const int j = jr + 16 * WD_NW * u + g4;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        m[u][s] = (j + s < jhi) ? (TR ? M[(int64_t)(j + s) * P.b + k0 + li] : mrow[j + s]) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < WD_NB; ++u) {
      const int j = jr + 16 * WD_NW * u + g4;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        v[u][s] = (j + s < jhi) ? sw_ld(src + (int64_t)(j + s) * WD_CW + c0 + li) : 0ull;
    }
#pragma unroll
    for (int u = 0; u < WD_NB; ++u) {
      const int j = jr + 16 * WD_NW * u + g4;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        if (v[u][s] == SW_EMPTY) {
          long long spins = 0;
          while ((v[u][s] = sw_ld(src + (int64_t)(j + s) * WD_CW + c0 + li)) == SW_EMPTY)
            if (!wd_spin(P, spins)) {
              v[u][s] = 0ull;
              break;
            }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < WD_NB; ++u)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        acc = mfma_f64(m[u][s], __longlong_as_double((long long)v[u][s]), acc);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
    red[wave * 256 + acc_row_f64(lane, r) * 16 + acc_col(lane)] = acc[r];
  __syncthreads();
  const int t = threadIdx.x;  // threads 0 .. 255: one output entry each
  if (t < 256) {
    double o = red[t];
#pragma unroll
    for (int w = 1; w < WD_NW; ++w) o += red[w * 256 + t];
    out[t] = o;
  }
  __syncthreads();
}

// Item of the k-th pick (w = blockIdx.x + k G) of a workgroup: workgroups
// are dealt to the 8 XCDs round-robin, so each XCD gets a contiguous range
// of items, i.e. whole row tiles (the four column tiles of 16 rows share
// their factor rows in that XCD's L2).  A bijection of [0, nitems).
__device__ __forceinline__ int wd_item(int w, int nitems) {
  const int q = nitems / 8, r = nitems % 8;
  const int x = w % 8, i = w / 8;
  return x * q + min(x, r) + i;
}

template <int TRANS>
__global__ void __launch_bounds__(WD_THREADS) pobtas_wide_kernel(WideSweepP P) {
  __shared__ double red[WD_NW * 256];
  __shared__ double out[256];
  const int b = P.b, n = P.n;
  const int ntc = (P.r + WD_T - 1) / WD_T;  // column tiles with at least one column
  const int nitems = (b / WD_T) * ntc;
  const int t = threadIdx.x, row = (t >> 4) & 15, col = t & 15;
  const int64_t bb = (int64_t)b * b;
  for (int step = 0; step < n; ++step) {
    const int i = TRANS ? n - 1 - step : step;
    const bool has_prev = TRANS ? (i + 1 < n) : (i > 0);
    // phase A: t_i = b_i - L_{i,i-1} y_{i-1}  /  s_i = y'_i - L_{i+1,i}^T x_{i+1}
    for (int w = blockIdx.x; w < nitems; w += P.G) {
      const int item = wd_item(w, nitems);
      const int k0 = (item / ntc) * WD_T, c0 = (item % ntc) * WD_T;
      if (has_prev) {
        const double* M = TRANS ? P.L + (int64_t)i * bb : P.L + (int64_t)(i - 1) * bb;
        const double* src = P.sbuf + (int64_t)(TRANS ? i + 1 : i - 1) * b * WD_CW;
        wd_tile<TRANS != 0>(P, M, k0, src, c0, 0, b, red, out);
      } else {
        if (t < 256) out[t] = 0.0;
        __syncthreads();
      }
      if (t < 256) {
        const int c = c0 + col;
        const int64_t grow = (int64_t)i * b + k0 + row;
        const double v = (c < P.r) ? P.B[grow * P.ldb + c] - out[t] : 0.0;
        sw_st(P.tbuf + grow * WD_CW + c, v);
      }
      __syncthreads();
    }
    // phase B: y_i = Linv_i t_i (rows up to the tile's)  /  x_i = Linv_i^T s_i (from it on)
    for (int w = blockIdx.x; w < nitems; w += P.G) {
      const int item = wd_item(w, nitems);
      const int k0 = (item / ntc) * WD_T, c0 = (item % ntc) * WD_T;
      const int jlo = TRANS ? k0 : 0, jhi = TRANS ? b : k0 + WD_T;
      wd_tile<TRANS != 0>(P, P.Linv + (int64_t)i * bb, k0, P.tbuf + (int64_t)i * b * WD_CW, c0, jlo,
                          jhi, red, out);
      if (t < 256) {
        const int c = c0 + col;
        const int64_t grow = (int64_t)i * b + k0 + row;
        if (c < P.r) P.B[grow * P.ldb + c] = out[t];
        sw_st(P.sbuf + grow * WD_CW + c, (c < P.r) ? out[t] : 0.0);
      }
      __syncthreads();
    }
  }
}

// Bt[m][c] = B0[m][c] - sum_{i < nblk} Pt[i][m][c] (fixed order), for the
// forward tip right-hand side; m < a, c < r (row strides ldb, r).
__global__ void wide_tip_reduce_kernel(double* out, const double* B0, int64_t ldb,
                                       const double* Pt, int a, int r, int nblk) {
  const int ar = a * r;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < ar; idx += gridDim.x * blockDim.x) {
    const int m = idx / r, c = idx % r;
    double v = B0[(int64_t)m * ldb + c];
    for (int i = 0; i < nblk; ++i) v -= Pt[(int64_t)i * ar + idx];
    out[idx] = v;
  }
}

// dst[m][c] = src[m][c] (row strides lds, ldd), m < a, c < r.
__global__ void wide_copy_rows_kernel(double* dst, int64_t ldd, const double* src, int64_t lds,
                                      int a, int r) {
  const int ar = a * r;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < ar; idx += gridDim.x * blockDim.x) {
    const int m = idx / r, c = idx % r;
    dst[(int64_t)m * ldd + c] = src[(int64_t)m * lds + c];
  }
}

}  // namespace rocbta
