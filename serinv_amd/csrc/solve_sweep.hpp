// Persistent triangular sweep with the BTA block Cholesky factor (pobtas) on
// gfx950: one CHAIN workgroup per chunk of right-hand-side columns walks the
// 64-row tiles in order, HELPER workgroups stream the rest of the factor.
//
// Reference recurrence: src/serinv/algs/pobtas.py:80-126 (forward,
// y = L^{-1} B) and :127-159 (backward, x = L^{-T} y).  It is evaluated on
// 64-row tiles in global tile coordinates: main tile g = i * nt + q is rows
// q*64 .. of block i (nt = ceil(b / 64)); tip tiles g = NT + t (NT = n * nt)
// are rows t*64 .. of the arrow tip.  L(r, j) is the factor tile of tile row
// r and tile column j; Linv(g) is the inverse of the diagonal tile, computed
// beforehand for every g.
//
// Forward (rows in increasing order):
//   chain step g:  y_g = Linv(g) (b_g - P_g - L(g, g-1) y_{g-1})
//   ROW(g)   (helper): P_g = sum_{j < g-1} L(g, j) y_j over the row band
//   ARR(i)   (helper): A_i = R_i y_i  (all arrow rows, block i)
//   tip step t:    y_t = Linv(NT+t) (b_t - sum_i A_i[t] - sum_{u<t} T(t,u) y_u)
// Backward (rows in decreasing order; tip first):
//   tip step t:    x_t = Linv(NT+t)^T (y_t - sum_{u>t} T(u,t)^T x_u)
//   chain step g:  x_g = Linv(g)^T (y_g - P_g - L(g+1, g)^T x_{g+1})
//   ROW(g)   (helper): P_g = sum_{j > g+1} L(j, g)^T x_j + sum_t R(t, g)^T x_t
//
// The chain keeps the previous tile's solution in LDS, so the only values it
// waits for are helper partial sums that were started many steps earlier.
// Helper tasks are handed out in chain order from one atomic queue
// (interleaved over the column chunks), so every task only waits on chain
// steps whose own inputs come from tasks that were handed out before it.
// Every wait is bounded; an overrun sets *status = -1 and all later waits
// return at once.  Results are published as data that is its own flag (see
// "dataflow" below).
#pragma once
#include "common.hpp"

namespace rocbta {

constexpr int SW_T = 64;          // tile rows
constexpr int SW_THREADS = 256;
constexpr int SW_LD = SW_T + 1;   // LDS pitch of a staged tile (doubles)

struct SweepP {
  const double* D;
  const double* L;
  const double* R;
  const double* T;
  double* B;
  int64_t ldb;
  int n, b, a, nt, ta, NT, NG;
  int r, rc, nchunks;  // columns, chunk width (template RC), chunks
  int trans;           // 0: forward, 1: backward
  const double* tinv;  // NG tiles of 64 x 64 (lower triangle valid)
  const double* tinv_tip;  // optional: the tip tiles' inverses (else tinv + NT tiles)
  double* ybuf;        // [chunk][NG][64 * RC] published solution tiles
  double* part;        // [chunk][NG][64 * RC] published row partial sums
  double* apart;       // [chunk][n][ta * 64 * RC] published arrow partial sums
  int* claim;
  int* counter;
  int* status;
  int ntasks;  // helper tasks, all chunks
  long long spin_limit;
};

// ------------------------------------------------------------ geometry
__host__ __device__ __forceinline__ const double* sw_tinv(const SweepP& P, int g) {
  if (g >= P.NT && P.tinv_tip) return P.tinv_tip + (int64_t)(g - P.NT) * SW_T * SW_T;
  return P.tinv + (int64_t)g * SW_T * SW_T;
}
__host__ __device__ __forceinline__ int sw_rows(const SweepP& P, int g) {
  if (g < P.NT) {
    const int q = g % P.nt;
    return (P.b - q * SW_T) < SW_T ? (P.b - q * SW_T) : SW_T;
  }
  const int t = g - P.NT;
  return (P.a - t * SW_T) < SW_T ? (P.a - t * SW_T) : SW_T;
}
__host__ __device__ __forceinline__ int64_t sw_row0(const SweepP& P, int g) {
  if (g < P.NT) return (int64_t)(g / P.nt) * P.b + (int64_t)(g % P.nt) * SW_T;
  return (int64_t)P.n * P.b + (int64_t)(g - P.NT) * SW_T;
}
// Factor tile (rr, j), rr >= j, rr in the band of j (or a tip row).
__host__ __device__ __forceinline__ const double* sw_tile(const SweepP& P, int rr, int j,
                                                         int64_t& ld) {
  if (rr >= P.NT) {
    const int t = rr - P.NT;
    if (j >= P.NT) {
      ld = P.a;
      return P.T + (int64_t)t * SW_T * P.a + (int64_t)(j - P.NT) * SW_T;
    }
    ld = P.b;
    return P.R + (int64_t)(j / P.nt) * P.a * P.b + (int64_t)t * SW_T * P.b +
           (int64_t)(j % P.nt) * SW_T;
  }
  const int jb = j / P.nt, jc = j % P.nt, rb = rr / P.nt, rc = rr % P.nt;
  ld = P.b;
  const double* base = (rb == jb) ? P.D : P.L;
  return base + (int64_t)jb * P.b * P.b + (int64_t)rc * SW_T * P.b + (int64_t)jc * SW_T;
}
// Exclusive end of the main band rows of tile column j.
__host__ __device__ __forceinline__ int sw_band_end(const SweepP& P, int j) {
  const int jb = j / P.nt;
  return (jb < P.n - 1) ? (jb + 2) * P.nt : (jb + 1) * P.nt;
}
// Helper ROW(g) contribution range [j0, j1) of main tiles.
__host__ __device__ __forceinline__ void sw_row_range(const SweepP& P, int g, int& j0, int& j1) {
  if (P.trans == 0) {
    const int ib = g / P.nt;
    j0 = (ib > 0) ? (ib - 1) * P.nt : 0;
    j1 = g - 1;
  } else {
    j0 = g + 2;
    j1 = sw_band_end(P, g);
  }
  if (j1 < j0) j1 = j0;
}
__host__ __device__ __forceinline__ bool sw_row_nonempty(const SweepP& P, int g) {
  int j0, j1;
  sw_row_range(P, g, j0, j1);
  return j1 > j0 || (P.trans == 1 && P.ta > 0);
}

// Helper task decoding.  Per chunk there are 2 * NT slots in chain order:
// forward slot 2s = ROW(s), 2s+1 = ARR(s / nt) when s is the last tile of its
// block; backward slot 2s = ROW(NT-1-s), 2s+1 unused.
enum SweepKind { SK_NOOP = 0, SK_ROW = 1, SK_ARR = 2 };
__host__ __device__ __forceinline__ int sw_decode(const SweepP& P, int task, int& chunk, int& g) {
  chunk = task % P.nchunks;
  const int local = task / P.nchunks;
  const int s = local >> 1, slot = local & 1;
  g = 0;
  if (s >= P.NT) return SK_NOOP;
  if (P.trans == 0) {
    if (slot == 0) {
      g = s;
      return sw_row_nonempty(P, g) ? SK_ROW : SK_NOOP;
    }
    if (P.ta > 0 && (s % P.nt) == P.nt - 1) {
      g = s / P.nt;  // block index
      return SK_ARR;
    }
    return SK_NOOP;
  }
  if (slot == 0) {
    g = P.NT - 1 - s;
    return sw_row_nonempty(P, g) ? SK_ROW : SK_NOOP;
  }
  return SK_NOOP;
}

// ------------------------------------------------------------ dataflow
// Every published block (solution tiles, partial sums) lives in a workspace
// buffer pre-filled with SW_EMPTY, a signalling-NaN bit pattern that
// arithmetic never produces.  Producers store the values with relaxed
// agent-scope atomic stores; consumers poll the values themselves with
// relaxed agent-scope atomic loads until none is SW_EMPTY.  The data is its
// own ready flag, so no fences are needed on either side.
constexpr unsigned long long SW_EMPTY = 0x7FF4DEADBEEF5A5Aull;

__device__ __forceinline__ unsigned long long sw_ld(const double* p) {
  return __hip_atomic_load(reinterpret_cast<unsigned long long*>(const_cast<double*>(p)),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void sw_st(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                     (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
// One more spin of a bounded wait; false once the wait must be abandoned
// (limit reached here, or another wave already reported a time-out).
__device__ __forceinline__ bool sw_spin(const SweepP& P, long long& spins) {
  if (++spins > P.spin_limit) {
    __hip_atomic_store(P.status, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
  }
  if ((spins & 63) == 0 && __hip_atomic_load(P.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0)
    return false;
  __builtin_amdgcn_s_sleep(1);
  return true;
}
// Value of a published element (0 after an abandoned wait).
__device__ __forceinline__ double sw_take1(const SweepP& P, const double* p) {
  unsigned long long v;
  long long spins = 0;
  while ((v = sw_ld(p)) == SW_EMPTY)
    if (!sw_spin(P, spins)) return 0.0;
  return __longlong_as_double((long long)v);
}
// Copy a published 64 x RC block into LDS.  Thread 0 waits for the first
// element (one poller per workgroup), then every thread takes its own.
// The caller synchronises before reading dst.
template <int RC>
__device__ __forceinline__ void sw_take(const SweepP& P, const double* src, double* dst) {
  if (threadIdx.x == 0) (void)sw_take1(P, src);
  __syncthreads();
  for (int idx = threadIdx.x; idx < SW_T * RC; idx += SW_THREADS) dst[idx] = sw_take1(P, src + idx);
}
template <int RC>
__device__ __forceinline__ void sw_put(double* dst, const double* src) {
  for (int idx = threadIdx.x; idx < SW_T * RC; idx += SW_THREADS) sw_st(dst + idx, src[idx]);
}

// ------------------------------------------------------------ data movement
typedef double sdbl2 __attribute__((ext_vector_type(2)));

// Thread t holds rows (t >> 3) + 32h (h = 0, 1), columns 8 (t & 7) .. + 7 of
// a 64 x 64 tile: one wave load covers 8 whole 512-byte rows.
struct SwRegs {
  double v[16];
};

template <bool LOWER>
__device__ __forceinline__ void sw_load(const double* p, int64_t ld, int rows, int cols,
                                        SwRegs& x) {
  const int r0 = threadIdx.x >> 3, c0 = (threadIdx.x & 7) * 8;
  const bool fast = !LOWER && rows == SW_T && cols == SW_T &&
                    ((reinterpret_cast<uintptr_t>(p) & 15) == 0) && ((ld & 1) == 0);
  if (fast) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const sdbl2 w = *reinterpret_cast<const sdbl2*>(
            __builtin_assume_aligned(p + (int64_t)(r0 + 32 * h) * ld + c0 + 2 * e, 16));
        x.v[8 * h + 2 * e] = w.x;
        x.v[8 * h + 2 * e + 1] = w.y;
      }
  } else {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int rr = r0 + 32 * h, cc = c0 + e;
        const bool in = rr < rows && cc < cols && (!LOWER || cc <= rr);
        x.v[8 * h + e] = in ? p[(int64_t)rr * ld + cc] : 0.0;
      }
  }
}
__device__ __forceinline__ void sw_store_lds(const SwRegs& x, double* S) {
  const int r0 = threadIdx.x >> 3, c0 = (threadIdx.x & 7) * 8;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int e = 0; e < 8; ++e) S[(r0 + 32 * h) * SW_LD + c0 + e] = x.v[8 * h + e];
}

// Vector block: rows [row0, row0 + rows) x chunk columns of B, zero padded
// to 64 x RC, element (k, c) at V[k * RC + c].
template <int RC>
struct SwVec {
  double v[(SW_T * RC + SW_THREADS - 1) / SW_THREADS];
};
template <int RC>
__device__ __forceinline__ void sw_vload(const SweepP& P, const double* Bp, int64_t ld, int rows,
                                         int cols, SwVec<RC>& x) {
  constexpr int PER = (SW_T * RC + SW_THREADS - 1) / SW_THREADS;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = threadIdx.x + e * SW_THREADS;
    const int k = idx / RC, c = idx % RC;
    x.v[e] = (idx < SW_T * RC && k < rows && c < cols) ? Bp[(int64_t)k * ld + c] : 0.0;
  }
}
template <int RC>
__device__ __forceinline__ void sw_vstore_lds(const SwVec<RC>& x, double* V) {
  constexpr int PER = (SW_T * RC + SW_THREADS - 1) / SW_THREADS;
#pragma unroll
  for (int e = 0; e < PER; ++e) {
    const int idx = threadIdx.x + e * SW_THREADS;
    if (idx < SW_T * RC) V[idx] = x.v[e];
  }
}

// acc[c] (+/-)= sum over this thread's quarter of the reduction index of
// A(o, k) V(k, c) (TR = false) or A(k, o) V(k, c) (TR = true), o = tid % 64.
template <int RC, bool TR, bool NEG>
__device__ __forceinline__ void sw_gemv(const double* A, const double* V, double (&acc)[RC]) {
  const int o = threadIdx.x & 63, kq = threadIdx.x >> 6;
#pragma unroll
  for (int kk = 0; kk < 16; ++kk) {
    const int k = kq * 16 + kk;
    const double av = TR ? A[k * SW_LD + o] : A[o * SW_LD + k];
#pragma unroll
    for (int c = 0; c < RC; ++c) acc[c] = NEG ? fma(-av, V[k * RC + c], acc[c]) : fma(av, V[k * RC + c], acc[c]);
  }
}
// out(o, c) = sum over the 4 quarters of acc (fixed order).
template <int RC>
__device__ __forceinline__ void sw_reduce(const double (&acc)[RC], double* red, double* out) {
  const int o = threadIdx.x & 63, kq = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < RC; ++c) red[(kq * SW_T + o) * RC + c] = acc[c];
  __syncthreads();
  for (int idx = threadIdx.x; idx < SW_T * RC; idx += SW_THREADS)
    out[idx] = ((red[idx] + red[SW_T * RC + idx]) + red[2 * SW_T * RC + idx]) + red[3 * SW_T * RC + idx];
  __syncthreads();
}

// ------------------------------------------------------------ roles
template <int RC>
struct SwLds {
  double A[SW_T * SW_LD];
  double red[4 * SW_T * RC];
  double V0[SW_T * RC];  // chain: previous solution; helper: staged vector
  double V1[SW_T * RC];  // chain: right-hand side / reduced vector
  double V2[SW_T * RC];  // chain: partial sum
};

// B rows of tile g, chunk columns.
template <int RC>
__device__ __forceinline__ double* sw_bptr(const SweepP& P, int chunk, int g) {
  return P.B + sw_row0(P, g) * P.ldb + (int64_t)chunk * RC;
}
template <int RC>
__device__ __forceinline__ double* sw_blk(double* base, const SweepP& P, int chunk, int g) {
  return base + ((int64_t)chunk * P.NG + g) * SW_T * RC;
}

// Write solution tile g (in LDS V) to B and publish it.
template <int RC>
__device__ __forceinline__ void sw_emit(const SweepP& P, int chunk, int g, int cols, const double* V) {
  const int m = sw_rows(P, g);
  double* bp = sw_bptr<RC>(P, chunk, g);
  for (int idx = threadIdx.x; idx < m * RC; idx += SW_THREADS) {
    const int k = idx / RC, c = idx % RC;
    if (c < cols) bp[(int64_t)k * P.ldb + c] = V[idx];
  }
  sw_put<RC>(sw_blk<RC>(P.ybuf, P, chunk, g), V);
}

template <int RC>
__device__ __forceinline__ void sw_tip_step(const SweepP& P, int chunk, int t, int cols, SwLds<RC>& S) {
  const int o = threadIdx.x & 63, kq = threadIdx.x >> 6;
  const bool fwd = P.trans == 0;
  const int g = P.NT + t;
  const int m = sw_rows(P, g);
  double acc[RC];
#pragma unroll
  for (int c = 0; c < RC; ++c) acc[c] = 0.0;
  if (fwd) {
    // arrow partial sums of every block (quarter kq takes i = kq mod 4)
    for (int i = kq; i < P.n; i += 4) {
      const double* ap = P.apart + ((int64_t)chunk * P.n + i) * P.ta * SW_T * RC +
                         (int64_t)t * SW_T * RC + o * RC;
#pragma unroll
      for (int c = 0; c < RC; ++c) acc[c] -= sw_take1(P, ap + c);
    }
  }
  const int u0 = fwd ? 0 : t + 1, u1 = fwd ? t : P.ta;
  for (int u = u0; u < u1; ++u) {
    const int gu = P.NT + u;
    int64_t ld;
    SwRegs x;
    if (fwd)
      sw_load<false>(sw_tile(P, g, gu, ld), ld, m, sw_rows(P, gu), x);
    else
      sw_load<false>(sw_tile(P, gu, g, ld), ld, sw_rows(P, gu), m, x);
    __syncthreads();  // previous product done with A / V0
    sw_take<RC>(P, sw_blk<RC>(P.ybuf, P, chunk, gu), S.V0);
    sw_store_lds(x, S.A);
    __syncthreads();
    if (fwd)
      sw_gemv<RC, false, true>(S.A, S.V0, acc);
    else
      sw_gemv<RC, true, true>(S.A, S.V0, acc);
  }
  // quarter 0 adds the right-hand side
  SwVec<RC> v;
  sw_vload<RC>(P, sw_bptr<RC>(P, chunk, g), P.ldb, m, cols, v);
  __syncthreads();
  sw_vstore_lds<RC>(v, S.V1);
  __syncthreads();
  if (kq == 0) {
#pragma unroll
    for (int c = 0; c < RC; ++c) acc[c] += S.V1[o * RC + c];
  }
  sw_reduce<RC>(acc, S.red, S.V1);  // V1 = reduced right-hand side
  SwRegs x;
  sw_load<true>(sw_tinv(P, g), SW_T, m, m, x);
  sw_store_lds(x, S.A);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < RC; ++c) acc[c] = 0.0;
  if (fwd)
    sw_gemv<RC, false, false>(S.A, S.V1, acc);
  else
    sw_gemv<RC, true, false>(S.A, S.V1, acc);
  sw_reduce<RC>(acc, S.red, S.V0);
  sw_emit<RC>(P, chunk, g, cols, S.V0);
}

template <int RC>
__device__ __forceinline__ void sw_chain(const SweepP& P, int chunk, SwLds<RC>& S) {
  const int o = threadIdx.x & 63, kq = threadIdx.x >> 6;
  const int cols = min(RC, P.r - chunk * RC);
  const bool fwd = P.trans == 0;

  if (!fwd)
    for (int t = P.ta - 1; t >= 0; --t) sw_tip_step<RC>(P, chunk, t, cols, S);

  // main tiles: V0 holds the previous solution tile (zero before the first)
  __syncthreads();
  for (int idx = threadIdx.x; idx < SW_T * RC; idx += SW_THREADS) S.V0[idx] = 0.0;
  SwRegs rs, rt;  // prefetched: coupling tile to the previous chain tile, Linv tile
  SwVec<RC> rb;   // prefetched right-hand side
  // Inputs of step g: the tiles and B rows are loaded into registers and the
  // helper partial sum into V2 one step ahead.
  auto prefetch = [&](int g) {
    const int m = sw_rows(P, g);
    int64_t ld;
    const int gp = fwd ? g - 1 : g + 1;
    const bool has = fwd ? (g >= 1) : (g + 1 < P.NT);
    if (has) {
      if (fwd)
        sw_load<false>(sw_tile(P, g, gp, ld), ld, m, sw_rows(P, gp), rs);
      else
        sw_load<false>(sw_tile(P, gp, g, ld), ld, sw_rows(P, gp), m, rs);
    }
    sw_load<true>(sw_tinv(P, g), SW_T, m, m, rt);
    sw_vload<RC>(P, sw_bptr<RC>(P, chunk, g), P.ldb, m, cols, rb);
    if (sw_row_nonempty(P, g)) sw_take<RC>(P, sw_blk<RC>(P.part, P, chunk, g), S.V2);
  };
  prefetch(fwd ? 0 : P.NT - 1);
  for (int s = 0; s < P.NT; ++s) {
    const int g = fwd ? s : P.NT - 1 - s;
    const bool has = fwd ? (g >= 1) : (g + 1 < P.NT);
    const bool part = sw_row_nonempty(P, g);
    __syncthreads();  // V2 filled; previous step done with A / V1
    if (has) sw_store_lds(rs, S.A);
    sw_vstore_lds<RC>(rb, S.V1);
    __syncthreads();
    double acc[RC];
#pragma unroll
    for (int c = 0; c < RC; ++c)
      acc[c] = (kq == 0) ? S.V1[o * RC + c] - (part ? S.V2[o * RC + c] : 0.0) : 0.0;
    if (has) {
      if (fwd)
        sw_gemv<RC, false, true>(S.A, S.V0, acc);
      else
        sw_gemv<RC, true, true>(S.A, S.V0, acc);
    }
    sw_reduce<RC>(acc, S.red, S.V1);  // V1 = reduced right-hand side (ends with a barrier)
    sw_store_lds(rt, S.A);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < RC; ++c) acc[c] = 0.0;
    if (fwd)
      sw_gemv<RC, false, false>(S.A, S.V1, acc);
    else
      sw_gemv<RC, true, false>(S.A, S.V1, acc);
    sw_reduce<RC>(acc, S.red, S.V0);  // V0 = solution tile g
    sw_emit<RC>(P, chunk, g, cols, S.V0);
    if (s + 1 < P.NT) prefetch(fwd ? g + 1 : g - 1);
  }
  if (fwd)
    for (int t = 0; t < P.ta; ++t) sw_tip_step<RC>(P, chunk, t, cols, S);
}

// Helper: accumulate the contributions of one task, in the order the chain
// publishes the solution tiles they need, with the next factor tile
// prefetched into registers.
template <int RC>
__device__ __forceinline__ void sw_helper(const SweepP& P, int kind, int chunk, int g, SwLds<RC>& S) {
  const int cols = min(RC, P.r - chunk * RC);
  (void)cols;
  const bool fwd = P.trans == 0;
  double acc[RC];
  if (kind == SK_ARR) {
    const int i = g;
    for (int t = 0; t < P.ta; ++t) {
#pragma unroll
      for (int c = 0; c < RC; ++c) acc[c] = 0.0;
      const int gt = P.NT + t;
      for (int q = 0; q < P.nt; ++q) {
        const int j = i * P.nt + q;
        int64_t ld;
        SwRegs x;
        sw_load<false>(sw_tile(P, gt, j, ld), ld, sw_rows(P, gt), sw_rows(P, j), x);
        __syncthreads();  // previous product done with A / V0
        sw_take<RC>(P, sw_blk<RC>(P.ybuf, P, chunk, j), S.V0);
        sw_store_lds(x, S.A);
        __syncthreads();
        sw_gemv<RC, false, false>(S.A, S.V0, acc);
      }
      sw_reduce<RC>(acc, S.red, S.V1);
      sw_put<RC>(P.apart + ((int64_t)chunk * P.n + i) * P.ta * SW_T * RC + (int64_t)t * SW_T * RC,
                 S.V1);
    }
    return;
  }
  // ROW(g): contribution k -> tile column / row jk
  int j0, j1;
  sw_row_range(P, g, j0, j1);
  const int nmain = j1 - j0;
  const int ntip = fwd ? 0 : P.ta;
  const int total = nmain + ntip;
  const int mg = sw_rows(P, g);
  // forward: j0 .. j1-1; backward: tip tiles ta-1 .. 0, then j1-1 .. j0
  auto jof = [&](int k) -> int {
    if (fwd) return j0 + k;
    if (k < ntip) return P.NT + (ntip - 1 - k);
    return j1 - 1 - (k - ntip);
  };
  auto load = [&](int jj, SwRegs& x) {
    int64_t ld;
    if (fwd)
      sw_load<false>(sw_tile(P, g, jj, ld), ld, mg, sw_rows(P, jj), x);
    else
      sw_load<false>(sw_tile(P, jj, g, ld), ld, sw_rows(P, jj), mg, x);
  };
#pragma unroll
  for (int c = 0; c < RC; ++c) acc[c] = 0.0;
  SwRegs x;
  load(jof(0), x);
  for (int k = 0; k < total; ++k) {
    const int jk = jof(k);
    __syncthreads();  // previous product done with A / V0
    sw_take<RC>(P, sw_blk<RC>(P.ybuf, P, chunk, jk), S.V0);
    sw_store_lds(x, S.A);
    __syncthreads();
    if (k + 1 < total) load(jof(k + 1), x);
    if (fwd)
      sw_gemv<RC, false, false>(S.A, S.V0, acc);
    else
      sw_gemv<RC, true, false>(S.A, S.V0, acc);
  }
  sw_reduce<RC>(acc, S.red, S.V1);
  sw_put<RC>(sw_blk<RC>(P.part, P, chunk, g), S.V1);
}

template <int RC>
__global__ void __launch_bounds__(SW_THREADS) pobtas_sweep_kernel(SweepP P) {
  __shared__ __attribute__((aligned(16))) SwLds<RC> S;
  __shared__ int s_task;
  if (threadIdx.x == 0) s_task = atomicAdd(P.claim, 1);
  __syncthreads();
  const int first = s_task;
  __syncthreads();
  if (first < P.nchunks) {
    sw_chain<RC>(P, first, S);
    return;
  }
  for (;;) {
    if (threadIdx.x == 0) s_task = atomicAdd(P.counter, 1);
    __syncthreads();
    const int task = s_task;
    __syncthreads();
    if (task >= P.ntasks) return;
    int chunk, g;
    const int kind = sw_decode(P, task, chunk, g);
    if (kind == SK_NOOP) continue;
    sw_helper<RC>(P, kind, chunk, g, S);
  }
}

// Many independent systems in one persistent launch (batched plans): claims
// below chain_begin[nmat] become the chains of (matrix, chunk) pairs
// (chain_begin: prefix sums of the matrices' chunk counts); every other
// workgroup serves the matrices' helper queues round-robin from each
// matrix's own in-order counter and leaves when a full round finds every
// queue drained.
template <int RC>
__global__ void __launch_bounds__(SW_THREADS)
    pobtas_sweep_multi_kernel(const SweepP* Ps, const int* chain_begin, int nmat, int* gclaim) {
  __shared__ __attribute__((aligned(16))) SwLds<RC> S;
  __shared__ int s_task;
  if (threadIdx.x == 0) s_task = atomicAdd(gclaim, 1);
  __syncthreads();
  const int first = s_task;
  __syncthreads();
  if (first < chain_begin[nmat]) {
    int m = 0;
    while (chain_begin[m + 1] <= first) ++m;
    sw_chain<RC>(Ps[m], first - chain_begin[m], S);
    return;
  }
  int m = (int)(blockIdx.x % (unsigned)nmat);
  int drained = 0;
  while (drained < nmat) {
    const SweepP& P = Ps[m];
    if (threadIdx.x == 0) s_task = atomicAdd(P.counter, 1);
    __syncthreads();
    const int task = s_task;
    __syncthreads();
    if (task >= P.ntasks) {
      ++drained;
    } else {
      drained = 0;
      int chunk, g;
      const int kind = sw_decode(P, task, chunk, g);
      if (kind != SK_NOOP) sw_helper<RC>(P, kind, chunk, g, S);
    }
    m = (m + 1 == nmat) ? 0 : m + 1;
  }
}

}  // namespace rocbta
