// Block-level persistent triangular sweep with the BTA block Cholesky factor
// (pobtas) on gfx950, for large diagonal blocks.
//
// Reference recurrence: src/serinv/algs/pobtas.py:80-126 (forward,
// y = L^{-1} B) and :127-159 (backward, x = L^{-T} y).  With the inverses
// Linv_i = L_ii^{-1} of the diagonal blocks computed beforehand (a batched
// trtri, independent over i), each block row is two dependent device-wide
// products instead of b / 64 dependent tile steps:
//
//   forward,  i = 0 .. n-1:   t_i = b_i - L_{i,i-1} y_{i-1}        (phase A)
//                             y_i = Linv_i t_i                      (phase B)
//             tip:            y_t = Tinv (b_t - sum_i R_i y_i)
//   backward, tip first:      x_t = Tinv^T y_t
//             i = n-1 .. 0:   s_i = y_i - L_{i+1,i}^T x_{i+1} - R_i^T x_t   (phase A)
//                             x_i = Linv_i^T s_i                            (phase B)
//
// `partial` (pobtas.py:99, 129) skips the last block's solve and the tip:
// forward, y_{n-1} = t_{n-1} and b_t -= sum_{i<n-1} R_i y_i is left unsolved;
// backward, x_{n-1} and x_t are the given B rows.
//
// The backward sweep reads the transposes of L_{i+1,i} and Linv_i, written
// beforehand (transpose_blocks_kernel), so that both directions read whole
// matrix rows: with the factor's own layout every output row of the
// backward products is a matrix column, read in 8-byte pieces.
//
// Work split: the b rows of a block are cut into row groups of RW rows;
// workgroup w owns row groups w, w + G, w + 2G, ... of every block (G
// workgroups, all co-resident).  A row group's phase A needs the whole
// previous solution block; its forward phase B needs t of the rows up to
// its own (a prefix), its backward phase B the rows from its own on (a
// suffix).  Inside a workgroup the reduction index is split over the 256
// threads and the partial sums are added in a fixed order, so results do not
// depend on scheduling.  Every published value lives in a buffer pre-filled
// with SW_EMPTY and is stored with a relaxed agent-scope atomic store;
// consumers poll the values themselves (the data is its own flag, as in
// solve_sweep.hpp).  Every wait is bounded; an overrun sets *status = -1 and
// later waits return at once.
#pragma once
#include "solve_sweep.hpp"

namespace rocbta {

constexpr int BS_THREADS = 256;
// Rows per row group (RW * CW partial sums per thread): four in both
// directions.
__host__ __device__ constexpr int bs_rw(int cw, int trans) { return 4 + 0 * cw * trans; }

// Transposes of nb square b x b blocks, 64 x 64 tiles through LDS (pitch 65,
// so that the column reads do not share banks).  Out of place: block z of
// src (stride sblk) -> block z of dst; workgroup (tile (I, J), z).
__global__ void __launch_bounds__(256) transpose_blocks_kernel(const double* src, double* dst,
                                                               int b, int64_t sblk, int T) {
  __shared__ double t[64][65];
  const int I = blockIdx.x % T, J = blockIdx.x / T;
  const int64_t off = (int64_t)blockIdx.y * sblk;
  const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = r0 + 4 * q, row = 64 * I + r, col = 64 * J + c;
    t[r][c] = (row < b && col < b) ? src[off + (int64_t)row * b + col] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = r0 + 4 * q, row = 64 * J + r, col = 64 * I + c;
    if (row < b && col < b) dst[off + (int64_t)row * b + col] = t[c][r];
  }
}

// The upper triangle of X^T for a lower-triangular X, in place: workgroup
// (pair p, z) writes tile (I, J) = tile (J, I)^T for I < J and transposes a
// diagonal tile within itself; the strictly lower tiles are left as they
// are (half the traffic of the full swap).  For readers of the upper
// triangle only (the backward block sweep reads Linv_i^T from the diagonal
// on); the strictly upper tiles need not hold anything beforehand.
__global__ void __launch_bounds__(256) transpose_lower_to_upper_kernel(double* X, int b,
                                                                       int64_t sblk) {
  __shared__ double t[64][65];
  int J = (int)((__builtin_sqrtf(8.0f * (float)blockIdx.x + 1.0f) - 1.0f) * 0.5f);
  while ((J + 1) * (J + 2) / 2 <= (int)blockIdx.x) ++J;
  while (J * (J + 1) / 2 > (int)blockIdx.x) --J;
  const int I = (int)blockIdx.x - J * (J + 1) / 2;
  double* x = X + (int64_t)blockIdx.y * sblk;
  const int c = threadIdx.x & 63, r0 = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = r0 + 4 * q;
    const int rb = 64 * J + r, cb = 64 * I + c;
    t[r][c] = (rb < b && cb < b) ? x[(int64_t)rb * b + cb] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int r = r0 + 4 * q;
    const int ra = 64 * I + r, ca = 64 * J + c;
    if (ra < b && ca < b) x[(int64_t)ra * b + ca] = t[c][r];  // tile (I, J) = tile (J, I)^T
  }
}

struct BlockSweepP {
  const double* L;     // (n-1, b, b); backward: the transposes L_{i+1,i}^T
  const double* R;     // (n, a, b)
  const double* Linv;  // (n, b, b): inverses of the diagonal blocks (dense, zero above);
                       // backward: their transposes (zero below)
  const double* Tinv;  // (a, a): inverse of the tip factor (zero above)
  double* B;           // right-hand sides, columns [0, r), row stride ldb
  int64_t ldb;
  int n, b, a, r;      // r <= CW (template)
  int trans, partial;
  int G;               // workgroups
  double* sbuf;        // (n*b, CW): published solution rows (y forward, x backward)
  double* tbuf;        // (n*b, CW): published phase-A rows (t forward, s backward)
  double* tpart;       // (G, a, CW): published per-workgroup tip partial sums (forward)
  double* tipbuf;      // (a, CW): published x_t (backward)
  int* status;
  long long spin_limit;
};

// Spin helpers on BlockSweepP (same protocol as sw_spin / sw_take1).
__device__ __forceinline__ bool bs_spin(const BlockSweepP& P, long long& spins) {
  if (++spins > P.spin_limit) {
    __hip_atomic_store(P.status, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
  }
  if ((spins & 63) == 0 && __hip_atomic_load(P.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0)
    return false;
  __builtin_amdgcn_s_sleep(1);
  return true;
}
__device__ __forceinline__ double bs_take(const BlockSweepP& P, const double* p) {
  unsigned long long v;
  long long spins = 0;
  while ((v = sw_ld(p)) == SW_EMPTY)
    if (!bs_spin(P, spins)) return 0.0;
  return __longlong_as_double((long long)v);
}

// out[v] = sum over the 256 threads of acc[v] (v < NV), fixed order: each
// wave reduces its 64 lanes through LDS, then the four wave sums are added
// in wave order.  red: 256 * NV doubles of LDS.  Ends with a barrier.
template <int NV>
__device__ __forceinline__ void bs_reduce(const double (&acc)[NV], double* red, double* out) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int v = 0; v < NV; ++v) red[v * BS_THREADS + tid] = acc[v];
  __syncthreads();
  // NV * 4 partial sums of 64 lanes each: thread (v, wave) sums one of them
  if (tid < NV * 4) {
    const int v = tid >> 2, wv = tid & 3;
    const double* p = red + v * BS_THREADS + wv * 64;
    double s = 0.0;
#pragma unroll 8
    for (int l = 0; l < 64; ++l) s += p[l];
    out[BS_THREADS + tid] = s;  // staging above the results
  }
  __syncthreads();
  if (tid < NV) {
    const double* q = out + BS_THREADS + tid * 4;
    out[tid] = ((q[0] + q[1]) + q[2]) + q[3];
  }
  __syncthreads();
}

// Tip entries (m, c) per thread: the block sweep takes a * CW <= BS_TIPMAX.
constexpr int BS_TIPMAX = 2 * BS_THREADS;

template <int CW, int RW>
struct BsLds {
  double red[RW * CW * BS_THREADS];
  double out[BS_THREADS + RW * CW * 4];
  double xt[BS_TIPMAX];  // x_t (backward) / the tip right-hand side (forward)
  double rv[BS_TIPMAX * RW];  // backward: R_i[m][k0 + q] at m * RW + q (a <= BS_TIPMAX)
};

// Rows of block i (k0 .. k0 + RW - 1) times the solution block `src`
// (forward: M[k][j] row access, backward: M[j][k] column access), K range
// [jlo, jhi), into acc[q][c].  src rows are polled from the buffer: a thread
// first issues the matrix loads and the polled loads of BS_KU reduction
// indices together, then re-polls only the values still marked empty, so
// one round trip covers all of them.
constexpr int BS_KU = 4;

template <int CW, int RW, bool TR>
__device__ __forceinline__ void bs_product(const BlockSweepP& P, const double* M, int k0,
                                           const double* src, int jlo, int jhi,
                                           double (&acc)[RW][CW]) {
  const int b = P.b;
  for (int j0 = jlo + (int)threadIdx.x; j0 < jhi; j0 += BS_KU * BS_THREADS) {
    double m[BS_KU][RW];
    unsigned long long v[BS_KU][CW];
#pragma unroll
    for (int u = 0; u < BS_KU; ++u) {
      const int j = j0 + u * BS_THREADS;
#pragma unroll
      for (int q = 0; q < RW; ++q) {
        const int k = k0 + q;
        m[u][q] = (j < jhi && k < b) ? (TR ? M[(int64_t)j * b + k] : M[(int64_t)k * b + j]) : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < BS_KU; ++u) {
      const int j = j0 + u * BS_THREADS;
#pragma unroll
      for (int c = 0; c < CW; ++c)
        v[u][c] = (j < jhi && c < P.r) ? sw_ld(src + (int64_t)j * CW + c) : 0ull;
    }
#pragma unroll
    for (int u = 0; u < BS_KU; ++u) {
      const int j = j0 + u * BS_THREADS;
#pragma unroll
      for (int c = 0; c < CW; ++c) {
        if (v[u][c] == SW_EMPTY) {
          long long spins = 0;
          while ((v[u][c] = sw_ld(src + (int64_t)j * CW + c)) == SW_EMPTY)
            if (!bs_spin(P, spins)) {
              v[u][c] = 0ull;
              break;
            }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < BS_KU; ++u)
#pragma unroll
      for (int q = 0; q < RW; ++q)
#pragma unroll
        for (int c = 0; c < CW; ++c)
          acc[q][c] = fma(m[u][q], __longlong_as_double((long long)v[u][c]), acc[q][c]);
  }
}

template <int CW, int RW>
__device__ __forceinline__ void bs_zero(double (&acc)[RW][CW]) {
#pragma unroll
  for (int q = 0; q < RW; ++q)
#pragma unroll
    for (int c = 0; c < CW; ++c) acc[q][c] = 0.0;
}

// Publish rows k0.. of block i (values in S.out[q * CW + c]) to buf, and
// optionally write them to B.
template <int CW, int RW>
__device__ __forceinline__ void bs_emit(const BlockSweepP& P, double* buf, int i, int k0,
                                        const double* vals, bool to_b) {
  const int tid = threadIdx.x;
  if (tid < RW * CW) {
    const int q = tid / CW, c = tid % CW, k = k0 + q;
    if (k < P.b && c < P.r) {
      const int64_t row = (int64_t)i * P.b + k;
      if (to_b) P.B[row * P.ldb + c] = vals[tid];
      sw_st(buf + row * CW + c, vals[tid]);
    }
  }
}

template <int CW, int TRANS>
__global__ void __launch_bounds__(BS_THREADS) pobtas_block_kernel(BlockSweepP P) {
  constexpr int RW = bs_rw(CW, TRANS);
  __shared__ BsLds<CW, RW> S;
  const int tid = threadIdx.x;
  const int b = P.b, n = P.n, a = P.a;
  const int ngroups = (b + RW - 1) / RW;
  double acc[RW][CW];
  if (TRANS == 0) {
    // ------------------------------------------------------------ forward
    // tip partial sums of this workgroup: thread tid owns entries
    // idx = tid + 256 u of the (m, c) pairs, idx < a * CW
    double tacc[BS_TIPMAX / BS_THREADS] = {};
    const int nsolve = P.partial ? n - 1 : n;
    for (int i = 0; i < n; ++i) {
      for (int g = blockIdx.x; g < ngroups; g += P.G) {
        const int k0 = g * RW;
        // phase A: t = b_i - L_{i,i-1} y_{i-1}
        bs_zero(acc);
        if (i > 0)
          bs_product<CW, RW, false>(P, P.L + (int64_t)(i - 1) * b * b, k0,
                                P.sbuf + (int64_t)(i - 1) * b * CW, 0, b, acc);
        double flat[RW * CW];
#pragma unroll
        for (int q = 0; q < RW; ++q)
#pragma unroll
          for (int c = 0; c < CW; ++c) flat[q * CW + c] = acc[q][c];
        bs_reduce<RW * CW>(flat, S.red, S.out);
        if (tid < RW * CW) {
          const int q = tid / CW, c = tid % CW, k = k0 + q;
          const double bv = (k < b && c < P.r) ? P.B[((int64_t)i * b + k) * P.ldb + c] : 0.0;
          S.out[tid] = bv - S.out[tid];
        }
        __syncthreads();
        // the last block of a partial solve keeps t as its result
        bs_emit<CW, RW>(P, P.tbuf, i, k0, S.out, i >= nsolve);
        __syncthreads();
      }
      if (i >= nsolve) continue;
      for (int g = blockIdx.x; g < ngroups; g += P.G) {
        const int k0 = g * RW;
        // phase B: y = Linv_i t (t rows 0 .. k0 + RW - 1)
        bs_zero(acc);
        bs_product<CW, RW, false>(P, P.Linv + (int64_t)i * b * b, k0, P.tbuf + (int64_t)i * b * CW, 0,
                              min(b, k0 + RW), acc);
        double flat[RW * CW];
#pragma unroll
        for (int q = 0; q < RW; ++q)
#pragma unroll
          for (int c = 0; c < CW; ++c) flat[q * CW + c] = acc[q][c];
        bs_reduce<RW * CW>(flat, S.red, S.out);
        bs_emit<CW, RW>(P, P.sbuf, i, k0, S.out, true);
        // tip: b_t -= R_i y_i over this row group
#pragma unroll
        for (int u = 0; u < BS_TIPMAX / BS_THREADS; ++u) {
          const int idx = tid + u * BS_THREADS;
          if (idx < a * CW) {
            const int m = idx / CW, c = idx % CW;
#pragma unroll
            for (int q = 0; q < RW; ++q) {
              const int k = k0 + q;
              if (k < b) tacc[u] = fma(P.R[((int64_t)i * a + m) * b + k], S.out[q * CW + c], tacc[u]);
            }
          }
        }
        __syncthreads();
      }
    }
    if (a == 0) return;
    // publish this workgroup's tip partial sums
#pragma unroll
    for (int u = 0; u < BS_TIPMAX / BS_THREADS; ++u) {
      const int idx = tid + u * BS_THREADS;
      if (idx < a * CW) sw_st(P.tpart + (int64_t)blockIdx.x * a * CW + idx, tacc[u]);
    }
    if (blockIdx.x != 0) return;
    // workgroup 0: y_t = Tinv (b_t - sum_w part_w)   (partial: b_t - sum only)
    double* tip = S.xt;
    for (int idx = tid; idx < a * CW; idx += BS_THREADS) {
      const int m = idx / CW, c = idx % CW;
      double s = 0.0;
      for (int w = 0; w < P.G; ++w) s += bs_take(P, P.tpart + (int64_t)w * a * CW + idx);
      const double bv = (c < P.r) ? P.B[((int64_t)n * b + m) * P.ldb + c] : 0.0;
      tip[idx] = bv - s;
    }
    __syncthreads();
    for (int idx = tid; idx < a * CW; idx += BS_THREADS) {
      const int m = idx / CW, c = idx % CW;
      double v = tip[idx];
      if (!P.partial) {
        v = 0.0;
        for (int mm = 0; mm <= m; ++mm) v = fma(P.Tinv[(int64_t)m * a + mm], tip[mm * CW + c], v);
      }
      if (c < P.r) P.B[((int64_t)n * b + m) * P.ldb + c] = v;
    }
    return;
  }
  // -------------------------------------------------------------- backward
  // x_t: solved by workgroup 0 and published (or the given B rows if
  // partial); every workgroup keeps it in LDS
  if (a > 0) {
    if (!P.partial) {
      if (blockIdx.x == 0) {
        for (int idx = tid; idx < a * CW; idx += BS_THREADS) {
          const int m = idx / CW, c = idx % CW;
          double v = 0.0;
          for (int mm = m; mm < a; ++mm)
            v = fma(P.Tinv[(int64_t)mm * a + m],
                    (c < P.r) ? P.B[((int64_t)n * b + mm) * P.ldb + c] : 0.0, v);
          S.xt[idx] = v;
        }
        __syncthreads();  // every read of the tip rows of B precedes the writes
        for (int idx = tid; idx < a * CW; idx += BS_THREADS) {
          const int m = idx / CW, c = idx % CW;
          if (c < P.r) P.B[((int64_t)n * b + m) * P.ldb + c] = S.xt[idx];
          sw_st(P.tipbuf + idx, S.xt[idx]);
        }
      } else {
        for (int idx = tid; idx < a * CW; idx += BS_THREADS) S.xt[idx] = bs_take(P, P.tipbuf + idx);
      }
    } else {
      for (int idx = tid; idx < a * CW; idx += BS_THREADS) {
        const int m = idx / CW, c = idx % CW;
        S.xt[idx] = (c < P.r) ? P.B[((int64_t)n * b + m) * P.ldb + c] : 0.0;
      }
    }
    __syncthreads();
  }
  for (int i = n - 1; i >= 0; --i) {
    const bool skip = P.partial && i == n - 1;  // x_{n-1} is the given B block
    for (int g = blockIdx.x; g < ngroups; g += P.G) {
      const int k0 = g * RW;
      if (skip) {
        if (tid < RW * CW) {
          const int q = tid / CW, c = tid % CW, k = k0 + q;
          S.out[tid] = (k < b && c < P.r) ? P.B[((int64_t)i * b + k) * P.ldb + c] : 0.0;
        }
        __syncthreads();
        bs_emit<CW, RW>(P, P.sbuf, i, k0, S.out, false);
        __syncthreads();
        continue;
      }
      // phase A: s = y_i - L_{i+1,i}^T x_{i+1} - R_i^T x_t (P.L holds L^T).
      // The R_i entries of this row group are fetched by all threads at
      // once into LDS (ordered by the barriers of bs_reduce), not one after
      // the other inside the m loop below.
      for (int idx = tid; idx < a * RW; idx += BS_THREADS) {
        const int m = idx / RW, q = idx % RW, k = k0 + q;
        S.rv[idx] = (k < b) ? P.R[((int64_t)i * a + m) * b + k] : 0.0;
      }
      bs_zero(acc);
      if (i + 1 < n)
        bs_product<CW, RW, false>(P, P.L + (int64_t)i * b * b, k0, P.sbuf + (int64_t)(i + 1) * b * CW,
                              0, b, acc);
      double flat[RW * CW];
#pragma unroll
      for (int q = 0; q < RW; ++q)
#pragma unroll
        for (int c = 0; c < CW; ++c) flat[q * CW + c] = acc[q][c];
      bs_reduce<RW * CW>(flat, S.red, S.out);
      if (tid < RW * CW) {
        const int q = tid / CW, c = tid % CW, k = k0 + q;
        double v = 0.0;
        if (k < b && c < P.r) {
          v = P.B[((int64_t)i * b + k) * P.ldb + c] - S.out[tid];
          for (int m = 0; m < a; ++m) v = fma(-S.rv[m * RW + q], S.xt[m * CW + c], v);
        }
        S.out[tid] = v;
      }
      __syncthreads();
      bs_emit<CW, RW>(P, P.tbuf, i, k0, S.out, false);
      __syncthreads();
    }
    if (skip) continue;
    for (int g = blockIdx.x; g < ngroups; g += P.G) {
      const int k0 = g * RW;
      // phase B: x = Linv_i^T s (s rows k0 .. b - 1)
      bs_zero(acc);
      bs_product<CW, RW, false>(P, P.Linv + (int64_t)i * b * b, k0, P.tbuf + (int64_t)i * b * CW, k0,
                            b, acc);
      double flat[RW * CW];
#pragma unroll
      for (int q = 0; q < RW; ++q)
#pragma unroll
        for (int c = 0; c < CW; ++c) flat[q * CW + c] = acc[q][c];
      bs_reduce<RW * CW>(flat, S.red, S.out);
      bs_emit<CW, RW>(P, P.sbuf, i, k0, S.out, true);
      __syncthreads();
    }
  }
}

}  // namespace rocbta
