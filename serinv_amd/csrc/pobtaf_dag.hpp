// Persistent tile-DAG block Cholesky of an SPD BTA matrix (pobtaf) on gfx950.
//
// The reference recurrence (src/serinv/algs/pobtaf.py:115-183) is evaluated
// tile by tile (64x64 tiles).  Every output tile of the factor is one task:
//
//   DIAG(i,c)   : tile (c,c) of block D_i  -> potrf (+ its inverse, kept in
//                 workspace for the solves of the tiles below it)
//   OFFD(i,r,c) : tile (r,c), r > c, of D_i        -> * inv(L_cc)^T
//   LOW(i,r,c)  : tile (r,c) of L_{i+1,i} (L[i])   -> * inv(L_cc)^T
//   ARR(i,c)    : tile c of the arrow block R_i     -> * inv(L_cc)^T
//   TIPP(i)     : P_i = R_i R_i^T (partial tip update, a x a)
//   TIP         : T - P_0 - P_1 - ... - P_{n-1} -> potrf
//
// Each tile task is left-looking: it starts from the input tile and
// subtracts the products of already-final factor tiles in increasing column
// order (first the L_{i,i-1} columns of the previous block row, which carry
// the Schur update D_i -= L_{i,i-1} L_{i,i-1}^T, then the columns of its own
// block column), waiting on each source tile's completion flag.
//
// Scheduling: workgroups take task indices from an atomic counter in a
// topological order (block column, then tile column, then DIAG, OFFD, LOW,
// ARR), so a task only ever waits on tasks with smaller indices, which have
// already been taken by running workgroups: the schedule cannot deadlock.
// Every wait is bounded; on overrun the kernel records info = -1 and moves
// on instead of hanging.
//
// Memory ordering: a finished tile is published with an agent-scope release
// (L2 write-back) before its flag store; a waiter polls with relaxed loads
// and issues one agent-scope acquire (L2 invalidate) before reading.
#pragma once
#include "common.hpp"
#include "tile_f64.hpp"

namespace rocbta {

constexpr int DAG_THREADS = 256;
constexpr int DT = 64;  // tile size

struct PobtafDag {
  double* D;
  double* L;
  double* R;
  double* T;
  int n, b, a, nt, ha;
  int tpb;     // tasks per block column
  int ntasks;  // n * tpb (+1 final tip task when a > 0)
  int* counter;
  int* flags;
  int* info;
  double* linv;  // n * nt * 64 * 64: inverse of every diagonal tile
  double* tipp;  // n * a * a: P_i
  long long spin_limit;
};

struct DagTile {
  const double* p;
  int64_t ld;
  int rows, cols;
};

__host__ __device__ __forceinline__ int dag_base(const PobtafDag& P, int c) {
  return c * (P.nt + P.ha + 1) + c * (P.nt - 1) - (c * (c - 1)) / 2;
}
__host__ __device__ __forceinline__ int t_diag(const PobtafDag& P, int i, int c) {
  return i * P.tpb + dag_base(P, c);
}
__host__ __device__ __forceinline__ int t_offd(const PobtafDag& P, int i, int r, int c) {
  return i * P.tpb + dag_base(P, c) + (r - c);
}
__host__ __device__ __forceinline__ int t_low(const PobtafDag& P, int i, int r, int c) {
  return i * P.tpb + dag_base(P, c) + (P.nt - c) + r;
}
__host__ __device__ __forceinline__ int t_arr(const PobtafDag& P, int i, int c) {
  return i * P.tpb + dag_base(P, c) + (2 * P.nt - c);
}
__host__ __device__ __forceinline__ int t_tipp(const PobtafDag& P, int i) {
  return i * P.tpb + dag_base(P, P.nt);
}

__host__ __device__ __forceinline__ int imin(int x, int y) { return x < y ? x : y; }
__host__ __device__ __forceinline__ DagTile tile_D(const PobtafDag& P, int i, int r, int c) {
  const int64_t b = P.b;
  return {P.D + i * b * b + (int64_t)r * DT * b + c * DT, b, imin(DT, P.b - r * DT),
          imin(DT, P.b - c * DT)};
}
__host__ __device__ __forceinline__ DagTile tile_L(const PobtafDag& P, int i, int r, int c) {
  const int64_t b = P.b;
  return {P.L + i * b * b + (int64_t)r * DT * b + c * DT, b, imin(DT, P.b - r * DT),
          imin(DT, P.b - c * DT)};
}
__host__ __device__ __forceinline__ DagTile tile_R(const PobtafDag& P, int i, int c) {
  const int64_t b = P.b;
  return {P.R + (int64_t)i * P.a * b + c * DT, b, P.a, imin(DT, P.b - c * DT)};
}

// Task kinds.  TIP is the single task n * tpb (only when a > 0).
enum DagKind { K_DTILE = 0, K_LOW = 1, K_ARR = 2, K_TIPP = 3, K_TIP = 4 };

// Decode a task index into (kind, block column i, tile row r, tile column c).
__host__ __device__ __forceinline__ void dag_decode(const PobtafDag& P, int task, int& kind,
                                                    int& i, int& r, int& c) {
  r = 0;
  c = 0;
  if (P.ha && task == P.n * P.tpb) {
    kind = K_TIP;
    i = P.n;
    return;
  }
  i = task / P.tpb;
  const int local = task - i * P.tpb;
  while (c < P.nt && dag_base(P, c + 1) <= local) ++c;
  if (c == P.nt) {
    kind = K_TIPP;
    return;
  }
  const int o = local - dag_base(P, c);
  if (o < P.nt - c) {
    kind = K_DTILE;
    r = c + o;
  } else if (o < 2 * P.nt - c) {
    kind = K_LOW;
    r = o - (P.nt - c);
  } else {
    kind = K_ARR;
  }
}

// ------------------------------------------------------------- flag protocol
__device__ __forceinline__ void dag_publish(const PobtafDag& P, int task) {
  __syncthreads();  // every wave's stores of the tile have completed
  if (threadIdx.x == 0)
    __hip_atomic_store(&P.flags[task], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// A spin overrun anywhere sets info = -1; every later wait then returns at
// once, so a scheduling fault costs one time-out, not one per task.
__device__ __forceinline__ bool dag_poll(const PobtafDag& P, int task, long long& spins) {
  while (__hip_atomic_load(&P.flags[task], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    if (++spins > P.spin_limit) return false;
    if ((spins & 255) == 0 &&
        __hip_atomic_load(P.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0)
      return false;
    __builtin_amdgcn_s_sleep(2);
  }
  return true;
}

// Wait until tasks ta and tb (tb < 0: none) are final; all threads return
// after an acquire.  On a spin overrun records info = -1 (no hang).
__device__ __forceinline__ void dag_wait2(const PobtafDag& P, int ta, int tb) {
  if (threadIdx.x == 0) {
    long long spins = 0;
    bool ok = dag_poll(P, ta, spins);
    if (ok && tb >= 0) ok = dag_poll(P, tb, spins);
    if (!ok) __hip_atomic_store(P.info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// ------------------------------------------------------------ tile movement
// Thread t holds row (t >> 2), columns 16*(t & 3) .. +15 of a 64x64 tile.
__device__ __forceinline__ void tile_load_regs(const DagTile& t, double (&v)[16]) {
  const int m = threadIdx.x >> 2, k0 = (threadIdx.x & 3) * 16;
  const double* src = t.p + (int64_t)m * t.ld + k0;
  const bool rok = m < t.rows;
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = (rok && k0 + q < t.cols) ? src[q] : 0.0;
}

// k-major staging: buf[k][m] = tile(m, k), so that MFMA operand fetches are
// contiguous 16-element row segments.
__device__ __forceinline__ void tile_store_kmajor(const double (&v)[16], double (*buf)[TT_LD]) {
  const int m = threadIdx.x >> 2, k0 = (threadIdx.x & 3) * 16;
#pragma unroll
  for (int q = 0; q < 16; ++q) buf[k0 + q][m] = v[q];
}

// Accumulator layout: wave w owns the 32x32 quadrant (wm, wn) =
// (32*(w>>1), 32*(w&1)); acc[i][j] is its 16x16 sub-tile (i, j).
__device__ __forceinline__ void acc_zero(f64x4 (&acc)[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
}

__device__ __forceinline__ void acc_load(const DagTile& t, f64x4 (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm + i * 16 + acc_row_f64(lane, r), col = wn + j * 16 + acc_col(lane);
        acc[i][j][r] = (row < t.rows && col < t.cols) ? t.p[(int64_t)row * t.ld + col] : 0.0;
      }
}

__device__ __forceinline__ void acc_store(double* p, int64_t ld, int rows, int cols,
                                          const f64x4 (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm + i * 16 + acc_row_f64(lane, r), col = wn + j * 16 + acc_col(lane);
        if (row < rows && col < cols) p[(int64_t)row * ld + col] = acc[i][j][r];
      }
}

// Row-major staging of the accumulator (S[row][col]).
__device__ __forceinline__ void acc_to_lds(const f64x4 (&acc)[2][2], double (*S)[TT_LD]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        S[wm + i * 16 + acc_row_f64(lane, r)][wn + j * 16 + acc_col(lane)] = acc[i][j][r];
}

// acc += sgn * A B^T with As[k][m] = A(m,k), Bs[k][n] = B(n,k).
// lower_only: skip the strictly-upper 32x32 quadrant (symmetric targets).
__device__ __forceinline__ void mfma_nt(const double (*As)[TT_LD], const double (*Bs)[TT_LD],
                                        f64x4 (&acc)[2][2], double sgn, bool lower_only) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  if (lower_only && wn > wm) return;
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int kk = 0; kk < DT; kk += 4) {
    const double a0 = sgn * As[kk + lk][wm + li];
    const double a1 = sgn * As[kk + lk][wm + 16 + li];
    const double b0 = Bs[kk + lk][wn + li];
    const double b1 = Bs[kk + lk][wn + 16 + li];
    acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
    acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
    acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
    acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
  }
}

// ---------------------------------------------------------- contributions
struct Contrib {
  DagTile A, B;
  int fa, fb;
};

// j-th contribution of a task (see the header comment); returns false past
// the end.  kind: 0 D_i tile (r, c) (DIAG when r == c), 1 L[i] tile, 2 arrow
// tile c of R_i, 3 TIPP(i).
__host__ __device__ __forceinline__ bool dag_contrib(const PobtafDag& P, int kind, int i, int r, int c,
                                            int j, Contrib& out) {
  const int nprev = (i > 0 && kind != 1 && kind != 3) ? P.nt : 0;
  if (kind == 3) {
    if (j >= P.nt) return false;
    out.A = tile_R(P, i, j);
    out.B = out.A;
    out.fa = t_arr(P, i, j);
    out.fb = -1;
    return true;
  }
  if (j < nprev) {  // previous block row: column k = j of L_{i,i-1} (and R_{i-1})
    const int k = j;
    out.B = tile_L(P, i - 1, c, k);
    out.fb = t_low(P, i - 1, c, k);
    if (kind == 0) {
      out.A = tile_L(P, i - 1, r, k);
      out.fa = t_low(P, i - 1, r, k);
    } else {
      out.A = tile_R(P, i - 1, k);
      out.fa = t_arr(P, i - 1, k);
    }
    return true;
  }
  const int jj = j - nprev;  // own block column, tile column jj < c
  if (jj >= c) return false;
  out.B = tile_D(P, i, c, jj);
  out.fb = t_offd(P, i, c, jj);
  if (kind == 0) {
    out.A = tile_D(P, i, r, jj);
    out.fa = t_offd(P, i, r, jj);
  } else if (kind == 1) {
    out.A = tile_L(P, i, r, jj);
    out.fa = t_low(P, i, r, jj);
  } else {
    out.A = tile_R(P, i, jj);
    out.fa = t_arr(P, i, jj);
  }
  return true;
}

__global__ void __launch_bounds__(DAG_THREADS) pobtaf_dag_kernel(PobtafDag P) {
  __shared__ double buf0[TT][TT_LD];
  __shared__ double buf1[TT][TT_LD];
  __shared__ double wsm[TT][WS_LD];
  __shared__ int s_task;
  const int tid = threadIdx.x;

  for (;;) {
    if (tid == 0) s_task = atomicAdd(P.counter, 1);
    __syncthreads();
    const int task = s_task;
    __syncthreads();
    if (task >= P.ntasks) return;

    int kind, i, r, c;
    dag_decode(P, task, kind, i, r, c);
    f64x4 acc[2][2];
    if (kind == K_TIP) {
      // ---------------- TIP: T - sum_i P_i, then potrf
      const DagTile tt = {P.T, P.a, P.a, P.a};
      acc_load(tt, acc);
      for (int ii = 0; ii < P.n; ++ii) {
        dag_wait2(P, t_tipp(P, ii), -1);
        f64x4 pi[2][2];
        acc_load({P.tipp + (int64_t)ii * P.a * P.a, P.a, P.a, P.a}, pi);
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) acc[x][y] -= pi[x][y];
      }
      acc_to_lds(acc, buf0);
      __syncthreads();
      lds_potrf(buf0, P.a, P.info, (int64_t)P.n * P.b);
      for (int idx = tid; idx < P.a * P.a; idx += DAG_THREADS) {
        const int rr = idx / P.a, cc = idx % P.a;
        P.T[idx] = (cc <= rr) ? buf0[rr][cc] : 0.0;
      }
      dag_publish(P, task);
      continue;
    }
    if (kind == 1 && i == P.n - 1) {  // no L_{n,n-1}
      dag_publish(P, task);
      continue;
    }

    // ---------------- initial tile
    DagTile out;
    if (kind == 0)
      out = tile_D(P, i, r, c);
    else if (kind == 1)
      out = tile_L(P, i, r, c);
    else if (kind == 2)
      out = tile_R(P, i, c);
    else
      out = {P.tipp + (int64_t)i * P.a * P.a, P.a, P.a, P.a};
    if (kind == 3)
      acc_zero(acc);
    else
      acc_load(out, acc);
    const double sgn = (kind == 3) ? 1.0 : -1.0;
    const bool diag = (kind == 0 && r == c) || kind == 3;

    // ---------------- left-looking accumulation, operands prefetched one
    // contribution ahead through registers
    Contrib cur;
    if (dag_contrib(P, kind, i, r, c, 0, cur)) {
      double va[16], vb[16];
      dag_wait2(P, cur.fa, cur.fb);
      tile_load_regs(cur.A, va);
      tile_load_regs(cur.B, vb);
      for (int j = 0;; ++j) {
        __syncthreads();
        tile_store_kmajor(va, buf0);
        tile_store_kmajor(vb, buf1);
        __syncthreads();
        Contrib nxt;
        const bool more = dag_contrib(P, kind, i, r, c, j + 1, nxt);
        if (more) {
          dag_wait2(P, nxt.fa, nxt.fb);
          tile_load_regs(nxt.A, va);
          tile_load_regs(nxt.B, vb);
        }
        mfma_nt(buf0, buf1, acc, sgn, diag);
        if (!more) break;
      }
    }
    __syncthreads();

    if (kind == 3) {  // TIPP(i): store P_i
      acc_store(P.tipp + (int64_t)i * P.a * P.a, P.a, P.a, P.a, acc);
      dag_publish(P, task);
      continue;
    }

    if (kind == 0 && r == c) {
      // ---------------- DIAG: potrf + inverse
      const int m = out.rows;
      acc_to_lds(acc, buf0);
      __syncthreads();
      lds_potrf(buf0, m, P.info, (int64_t)i * P.b + (int64_t)c * DT);
      double* dst = const_cast<double*>(out.p);
      for (int idx = tid; idx < DT * DT; idx += DAG_THREADS) {
        const int rr = idx >> 6, cc = idx & 63;
        if (rr < m && cc < m) dst[(int64_t)rr * out.ld + cc] = (cc <= rr) ? buf0[rr][cc] : 0.0;
      }
      // zero the strict-upper tiles of this tile row
      const int zc = P.b - (c + 1) * DT;
      for (int rr = 0; rr < m; ++rr)
        for (int cc = tid; cc < zc; cc += DAG_THREADS) dst[(int64_t)rr * out.ld + m + cc] = 0.0;
      lds_trtri(buf0, buf1, wsm, m);
      double* li = P.linv + ((int64_t)i * P.nt + c) * DT * DT;
      for (int idx = tid; idx < DT * DT; idx += DAG_THREADS) li[idx] = buf1[idx >> 6][idx & 63];
      dag_publish(P, task);
      continue;
    }

    // ---------------- solve: out = acc * inv(L_cc)^T
    dag_wait2(P, t_diag(P, i, c), -1);
    {
      double vl[16];
      tile_load_regs({P.linv + ((int64_t)i * P.nt + c) * DT * DT, DT, DT, DT}, vl);
      // A operand = acc, k-major: buf0[k][m] = acc(m, k)
      const int lane = tid & 63, wave = tid >> 6;
      const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
          for (int q = 0; q < 4; ++q)
            buf0[wn + y * 16 + acc_col(lane)][wm + x * 16 + acc_row_f64(lane, q)] = acc[x][y][q];
      tile_store_kmajor(vl, buf1);
      __syncthreads();
      acc_zero(acc);
      mfma_nt(buf0, buf1, acc, 1.0, false);
      acc_store(const_cast<double*>(out.p), out.ld, out.rows, out.cols, acc);
    }
    dag_publish(P, task);
  }
}

}  // namespace rocbta
