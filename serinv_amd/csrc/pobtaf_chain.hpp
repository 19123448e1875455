// Persistent block Cholesky of an SPD BTA matrix (pobtaf) on gfx950:
// one CHAIN workgroup walks the diagonal tiles, HELPER workgroups do the rest.
//
// The reference recurrence (src/serinv/algs/pobtaf.py:115-183) is evaluated on
// 64x64 tiles.  Tiles are addressed in global tile coordinates: tile row /
// column g = i * nt + c is tile c of block i (nt = ceil(b / 64)); the band of
// tile column j (block jb) holds rows j .. (jb + 2) * nt - 1 (the rest of D_jb
// and all of L_{jb+1,jb}), plus the arrow tile of R_jb.  L(r, k) denotes a
// finished factor tile.
//
// Chain workgroup, step g (strictly sequential):
//   D(g,g) <- PD(g) - L(g,g-1) L(g,g-1)^T       (L(g,g-1) kept in LDS)
//   L(g,g) = chol(D(g,g)), Linv(g) = L(g,g)^{-1}  -> published
//   L(g+1,g) = PX(g) Linv(g)^T                    -> published, kept in LDS
// where PD(g) = D(g,g) minus the products of all columns k < g-1 and
// PX(g) = X(g+1,g) minus the products of all columns k < g are prepared by
// helpers.  Helper tasks:
//   TRSM(r, j), r >= j + 2 : X(r,j) - sum_k L(r,k) L(j,k)^T, then * Linv(j)^T
//   ARR(j)                 : the same for the arrow tile of column j
//   PX(g), PD(g)           : the partial sums above (stored in place)
//   TIPP(i)                : P_i = R_i R_i^T;  TIP: chol(T - P_0 - ... - P_{n-1})
// Helper tasks are taken from an atomic queue organised in stages; stage h
// holds the tasks that need chain outputs of steps <= h - 1 only, and inside
// a stage every task only waits on earlier slots.  The chain step g needs
// PD(g) (stage g-1) and PX(g) (stage g), so no wait can cycle.  Every wait is
// bounded: an overrun sets info = -1 and makes all later waits return.
//
// Memory ordering: tiles are published with an agent-scope release before
// the flag store; waiters poll relaxed and issue one agent-scope acquire.
#pragma once
#include "common.hpp"
#include "tile_f64.hpp"

namespace rocbta {

constexpr int CT = 64;          // tile size
constexpr int CH_THREADS = 256;
constexpr int KP = CT + 2;      // pitch of K-contiguous LDS tiles (doubles)

struct PobtafChain {
  double* D;
  double* L;
  double* R;
  double* T;
  int n, b, a, nt, ha;
  int NT;       // number of diagonal tiles (n * nt)
  int slots;    // helper slots per stage
  int nstages;  // NT + 2
  int ntasks;   // nstages * slots
  int* counter;  // helper queue head
  int* claim;    // chain claim ticket
  int* hflags;   // helper task flags
  int* cflags;   // chain flags: [2g] Linv(g) ready, [2g+1] L(g+1,g) ready
  int* info;
  int* chain_cu;  // 1 + hardware CU key of the chain workgroup (0 until known)
  double* linv;   // NT * 64 * 64
  double* tipp;   // n * a * a
  double* tip_inv;  // optional (batched plans): L_T^{-1} as a 64 x 64 tile
  long long spin_limit;
  // Mixed precision: the tile products (trsm, syrk/gemm updates) round their
  // FP64 operands to FP32 and accumulate in FP32 MFMA; the tiles stay FP64
  // in memory and the diagonal potrf / trtri and the tip run in FP64.
  int f32;
  // Chain schedule for full 64-wide tiles (b % 64 == 0): the sub-diagonal
  // solve is formed by column blocks around the tile factorization
  // (chain_run_fused); 0 selects the step-after-step schedule (chain_run).
  int fused;
  // Diagnostic timeline (nullptr normally): 100 MHz timestamps, CHAIN_TRACE_EV
  // per chain step g at trace[CHAIN_TRACE_EV * g + e], see chain_trace.
  long long* trace;
};

// Timeline events of the fused chain step g (diagnostic builds of a run):
//   0 step start, 1 factorization done, 2 PX(g)/PD(g+1) staged,
//   3 column blocks 0-2 and Linv(g) done, 4 L(g+1,g) complete, 5 step end,
//   6 PX(g) published (by its helper), 7 L(g+2,g) published (by its helper).
constexpr int CHAIN_TRACE_EV = 8;
__device__ __forceinline__ void chain_trace(const PobtafChain& P, int g, int e) {
  if (P.trace && threadIdx.x == 0)
    P.trace[CHAIN_TRACE_EV * g + e] = (long long)__builtin_amdgcn_s_memrealtime();
}

// Hardware key of the CU running the calling wave: XCC id and the
// SE/SH/CU fields of HW_ID.
__device__ __forceinline__ int hw_cu_key() {
  const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | (0 << 6) | 4);     // HW_ID
  const unsigned xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);    // XCC_ID
  return (int)(((xcc & 15u) << 8) | ((hw >> 8) & 0xffu));
}

struct CTile {
  double* p;
  int64_t ld;
  int rows, cols;
};

enum ChainKind { HK_NOOP = 0, HK_TRSM = 1, HK_PX = 2, HK_PD = 3, HK_ARR = 4, HK_TIPP = 5, HK_TIP = 6 };

__host__ __device__ __forceinline__ int cmin(int x, int y) { return x < y ? x : y; }
__host__ __device__ __forceinline__ int cmax(int x, int y) { return x > y ? x : y; }

// exclusive end of the band rows of tile column j
__host__ __device__ __forceinline__ int band_end(const PobtafChain& P, int j) {
  const int jb = j / P.nt;
  return (jb < P.n - 1) ? (jb + 2) * P.nt : (jb + 1) * P.nt;
}
// first contributing column for a row of block rb
__host__ __device__ __forceinline__ int kmin_of(const PobtafChain& P, int rb) {
  return cmax(0, (rb - 1) * P.nt);
}

__host__ __device__ __forceinline__ CTile tile_g(const PobtafChain& P, int r, int j) {
  const int jb = j / P.nt, jc = j % P.nt, rb = r / P.nt, rc = r % P.nt;
  const int64_t b = P.b;
  double* base = (rb == jb) ? P.D : P.L;
  return {base + (int64_t)jb * b * b + (int64_t)rc * CT * b + (int64_t)jc * CT, b,
          cmin(CT, P.b - rc * CT), cmin(CT, P.b - jc * CT)};
}
// Strict-upper D tile (jc, rc) that mirrors the lower tile (r, j) of the same
// block; the factor has zeros there.
__host__ __device__ __forceinline__ CTile tile_mirror(const PobtafChain& P, int r, int j) {
  const int jb = j / P.nt, jc = j % P.nt, rc = r % P.nt;
  const int64_t b = P.b;
  return {P.D + (int64_t)jb * b * b + (int64_t)jc * CT * b + (int64_t)rc * CT, b,
          cmin(CT, P.b - jc * CT), cmin(CT, P.b - rc * CT)};
}
__host__ __device__ __forceinline__ CTile tile_arrow(const PobtafChain& P, int j) {
  const int jb = j / P.nt, jc = j % P.nt;
  const int64_t b = P.b;
  return {P.R + (int64_t)jb * P.a * b + (int64_t)jc * CT, b, P.a, cmin(CT, P.b - jc * CT)};
}

// ---------------------------------------------------------- task indexing
// slot layout of stage h (j = h - 1):
//   0             TRSM(j+2, j)
//   1             PX(h)
//   2             PD(h+1)
//   3 .. 3+2nt-1  TRSM(j+3+q, j)
//   3+2nt         ARR(j)
//   4+2nt         TIPP(h/nt - 1) when h is a multiple of nt
// stage NT+1, slot 0: TIP
__host__ __device__ __forceinline__ int t_trsm(const PobtafChain& P, int r, int j) {
  const int h = j + 1;
  return h * P.slots + ((r == j + 2) ? 0 : 3 + (r - (j + 3)));
}
__host__ __device__ __forceinline__ int t_px(const PobtafChain& P, int g) { return g * P.slots + 1; }
__host__ __device__ __forceinline__ int t_pd(const PobtafChain& P, int g) {
  return (g - 1) * P.slots + 2;
}
__host__ __device__ __forceinline__ int t_arr(const PobtafChain& P, int j) {
  return (j + 1) * P.slots + 3 + 2 * P.nt;
}
__host__ __device__ __forceinline__ int t_tippc(const PobtafChain& P, int i) {
  return ((i + 1) * P.nt) * P.slots + 4 + 2 * P.nt;
}
__host__ __device__ __forceinline__ int t_tipc(const PobtafChain& P) {
  return (P.NT + 1) * P.slots;
}

// Decode a helper task: kind, and (r, j) = output tile (TRSM), g in j (PX/PD),
// column j (ARR), block i in j (TIPP).
__host__ __device__ __forceinline__ int chain_decode(const PobtafChain& P, int task, int& r,
                                                     int& j) {
  const int h = task / P.slots, slot = task % P.slots;
  r = 0;
  j = 0;
  if (h == P.NT + 1) return (slot == 0 && P.ha) ? HK_TIP : HK_NOOP;
  const int jj = h - 1;
  if (slot == 0) {
    if (jj >= 0 && jj + 2 < band_end(P, jj)) {
      r = jj + 2;
      j = jj;
      return HK_TRSM;
    }
    return HK_NOOP;
  }
  if (slot == 1) {
    if (h < P.NT && h + 1 < band_end(P, h)) {
      j = h;
      return HK_PX;
    }
    return HK_NOOP;
  }
  if (slot == 2) {
    if (h + 1 < P.NT) {
      j = h + 1;
      return HK_PD;
    }
    return HK_NOOP;
  }
  if (slot < 3 + 2 * P.nt) {
    const int rr = jj + 3 + (slot - 3);
    if (jj >= 0 && rr < band_end(P, jj)) {
      r = rr;
      j = jj;
      return HK_TRSM;
    }
    return HK_NOOP;
  }
  if (slot == 3 + 2 * P.nt) {
    if (P.ha && jj >= 0) {
      j = jj;
      return HK_ARR;
    }
    return HK_NOOP;
  }
  if (slot == 4 + 2 * P.nt && P.ha && h >= P.nt && h % P.nt == 0) {
    j = h / P.nt - 1;
    return HK_TIPP;
  }
  return HK_NOOP;
}

// Readiness flag of the factor tile (r, k), r > k (arrow: r < 0).
__host__ __device__ __forceinline__ int* factor_flag(const PobtafChain& P, int r, int k) {
  if (r < 0) return P.hflags + t_arr(P, k);
  if (r == k + 1) return P.cflags + 2 * k + 1;
  return P.hflags + t_trsm(P, r, k);
}

// Contribution list of a helper task: columns k in [k0, k1) with operands
// A = factor tile (ra, k) (ra < 0: arrow tile k) and B = factor tile (jb, k).
struct ContribRange {
  int ra, jb, k0, k1;
};
__host__ __device__ __forceinline__ ContribRange chain_contribs(const PobtafChain& P, int kind,
                                                               int r, int j) {
  switch (kind) {
    case HK_TRSM:
      return {r, j, kmin_of(P, r / P.nt), j};
    case HK_ARR:
      return {-1, j, kmin_of(P, j / P.nt), j};
    case HK_PX:
      return {j + 1, j, kmin_of(P, (j + 1) / P.nt), j};
    case HK_PD:
      return {j, j, kmin_of(P, j / P.nt), j - 1};
    default:
      return {0, 0, 0, 0};
  }
}

// ---------------------------------------------------------------- flags
__device__ __forceinline__ bool chain_poll(const PobtafChain& P, const int* f, long long& spins) {
  while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    if (++spins > P.spin_limit) return false;
    if ((spins & 255) == 0 &&
        __hip_atomic_load(P.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 0)
      return false;
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// Wait for flags fa (and fb if non-null); all threads return after an acquire.
__device__ __forceinline__ void chain_wait(const PobtafChain& P, const int* fa, const int* fb) {
  if (threadIdx.x == 0) {
    long long spins = 0;
    bool ok = chain_poll(P, fa, spins);
    if (ok && fb) ok = chain_poll(P, fb, spins);
    if (!ok) __hip_atomic_store(P.info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// The calling wave alone waits for fa (and fb if non-null), then acquires.
// All lanes poll the same address, so the loop is uniform; no barrier.
__device__ __forceinline__ void wave_wait(const PobtafChain& P, const int* fa, const int* fb) {
  long long spins = 0;
  bool ok = chain_poll(P, fa, spins);
  if (ok && fb) ok = chain_poll(P, fb, spins);
  if (!ok && (threadIdx.x & 63) == 0)
    __hip_atomic_store(P.info, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Every wave waits for its own global stores to complete (s_waitcnt
// vmcnt(0)) before the barrier: a workgroup barrier does not wait for them,
// and the release by thread 0 (L2 write-back, then its own wave's counter)
// would otherwise miss stores of other waves still on their way to L2.
// With all of them in L2, thread 0's single release covers the tile.
__device__ __forceinline__ void wave_stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// The explicit wait after thread 0's release fence keeps the flag store
// behind the completion of the L2 write-back: with the counters already
// drained, the compiler otherwise emits no wait between the two.
__device__ __forceinline__ void chain_publish(int* f, int publisher = 0) {
  wave_stores_done();
  __syncthreads();
  if ((int)threadIdx.x == publisher) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    wave_stores_done();
    __hip_atomic_store(f, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ----------------------------------------------------------- tile movement
typedef double cdbl2 __attribute__((ext_vector_type(2)));

// Thread t holds rows (t >> 5) + 8q (q = 0..7), columns 2*(t & 31) .. +1 of a
// 64x64 tile: each wave load instruction covers two whole 512-byte rows.
struct TileRegs16 {
  cdbl2 v[8];
};

__device__ __forceinline__ void tile_load16(const CTile& t, TileRegs16& x) {
  const int rs = threadIdx.x >> 5, c = (threadIdx.x & 31) * 2;
  const bool fast = t.rows == CT && t.cols == CT &&
                    ((reinterpret_cast<uintptr_t>(t.p) & 15) == 0) && ((t.ld & 1) == 0);
  if (fast) {
#pragma unroll
    for (int q = 0; q < 8; ++q)
      x.v[q] = *reinterpret_cast<const cdbl2*>(
          __builtin_assume_aligned(t.p + (int64_t)(rs + 8 * q) * t.ld + c, 16));
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int row = rs + 8 * q;
      const double* src = t.p + (int64_t)row * t.ld + c;
      x.v[q].x = (row < t.rows && c < t.cols) ? src[0] : 0.0;
      x.v[q].y = (row < t.rows && c + 1 < t.cols) ? src[1] : 0.0;
    }
  }
}

// K-contiguous staging: X[row][k] with pitch KP (row = m or n index).
__device__ __forceinline__ void tile_store16(const TileRegs16& x, double* X) {
  const int rs = threadIdx.x >> 5, c = (threadIdx.x & 31) * 2;
#pragma unroll
  for (int q = 0; q < 8; ++q) *reinterpret_cast<cdbl2*>(X + (rs + 8 * q) * KP + c) = x.v[q];
}

// acc (+)= sgn * A B^T over k in [0, 64), A = XA[m][k], B = XB[n][k].
template <typename ACC>
__device__ __forceinline__ void mfma_abt(const double* XA, const double* XB, ACC (&acc)[2][2],
                                         double sgn, bool lower_only) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  if (lower_only && wn > wm) return;
  const int li = lane & 15, lk = lane >> 4;
  // operands double-buffered across k steps: the LDS reads of step kk + 4
  // are issued before the MFMAs of step kk
  double a[2][2], b[2][2];
  auto fetch = [&](int buf, int kk) {
    a[buf][0] = XA[(wm + li) * KP + kk + lk];
    a[buf][1] = XA[(wm + 16 + li) * KP + kk + lk];
    b[buf][0] = XB[(wn + li) * KP + kk + lk];
    b[buf][1] = XB[(wn + 16 + li) * KP + kk + lk];
  };
  fetch(0, 0);
#pragma unroll
  for (int kk = 0; kk < CT; kk += 4) {
    const int cur = (kk >> 2) & 1;
    if (kk + 4 < CT) fetch(cur ^ 1, kk + 4);
    __builtin_amdgcn_sched_barrier(0);
    const double a0 = sgn * a[cur][0], a1 = sgn * a[cur][1];
    acc[0][0] = mfma_acc(a0, b[cur][0], acc[0][0]);
    acc[0][1] = mfma_acc(a0, b[cur][1], acc[0][1]);
    acc[1][0] = mfma_acc(a1, b[cur][0], acc[1][0]);
    acc[1][1] = mfma_acc(a1, b[cur][1], acc[1][1]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// acc += a32 (an FP32 accumulator in the FP32 MFMA register layout),
// remapped through the 64 x 64 LDS scratch X (row pitch `pitch`).  X must be
// free once every wave arrives; the call starts and ends with a barrier.
__device__ __forceinline__ void acc32_add(const f32x4 (&a32)[2][2], f64x4 (&acc)[2][2], double* X,
                                          int pitch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        X[(wm + i * 16 + acc_row_f32(lane, r)) * pitch + wn + j * 16 + acc_col(lane)] =
            (double)a32[i][j][r];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        acc[i][j][r] += X[(wm + i * 16 + acc_row_f64(lane, r)) * pitch + wn + j * 16 + acc_col(lane)];
  __syncthreads();
}

__device__ __forceinline__ void cacc32_zero(f32x4 (&acc)[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
}

__device__ __forceinline__ void cacc_zero(f64x4 (&acc)[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
}

__device__ __forceinline__ void cacc_load(const CTile& t, f64x4 (&acc)[2][2]) {
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

__device__ __forceinline__ void cacc_store(const CTile& t, const f64x4 (&acc)[2][2]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm + i * 16 + acc_row_f64(lane, r), col = wn + j * 16 + acc_col(lane);
        if (row < t.rows && col < t.cols) t.p[(int64_t)row * t.ld + col] = acc[i][j][r];
      }
}

// Stage the accumulator in LDS: X[row * pitch + col] = acc(row, col).
__device__ __forceinline__ void cacc_to_lds(const f64x4 (&acc)[2][2], double* X, int pitch) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        X[(wm + i * 16 + acc_row_f64(lane, r)) * pitch + wn + j * 16 + acc_col(lane)] =
            acc[i][j][r];
}

// Store a tile staged in LDS (X[row * pitch + col]) to memory.  Full,
// 16-byte aligned tiles: each thread reads its 8 row pairs (rows rs + 8q,
// columns c, c + 1) first and then issues 16-byte stores; other tiles go
// element by element.
__device__ __forceinline__ void lds_tile_store(const double* X, int pitch, const CTile& t) {
  const int tid = threadIdx.x;
  if (t.rows == CT && t.cols == CT && (t.ld & 1) == 0 &&
      (reinterpret_cast<uintptr_t>(t.p) & 15) == 0) {
    const int rs = tid >> 5, c = (tid & 31) * 2;
    cdbl2 v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      v[q].x = X[(rs + 8 * q) * pitch + c];
      v[q].y = X[(rs + 8 * q) * pitch + c + 1];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) *reinterpret_cast<cdbl2*>(t.p + (int64_t)(rs + 8 * q) * t.ld + c) = v[q];
  } else {
    for (int idx = tid; idx < CT * CT; idx += CH_THREADS) {
      const int row = idx >> 6, col = idx & 63;
      if (row < t.rows && col < t.cols) t.p[(int64_t)row * t.ld + col] = X[row * pitch + col];
    }
  }
}

// Zero a whole tile (used for the strict-upper tiles of the D blocks).
__device__ __forceinline__ void tile_zero(const CTile& t) {
  for (int idx = threadIdx.x; idx < CT * CT; idx += CH_THREADS) {
    const int row = idx >> 6, col = idx & 63;
    if (row < t.rows && col < t.cols) t.p[(int64_t)row * t.ld + col] = 0.0;
  }
}

// ------------------------------------------------------------ helper work
// acc -= sum over the contribution range (A = factor tile (ra,k) or arrow
// tile k, B = factor tile (jb,k)); operands prefetched one product ahead.
// TAG gives each kernel its own copy of the non-inlined workers, so that
// each copy sees a single kind of descriptor pointer (kernel argument or
// global array) and is optimised for it.
template <typename ACC>
__device__ __forceinline__ void helper_products(const PobtafChain& P, const ContribRange& cr,
                                                double sgn, bool lower_only, double* XA,
                                                double* XB, ACC (&acc)[2][2]) {
  TileRegs16 ra, rb;
  auto tileA = [&](int k) { return cr.ra < 0 ? tile_arrow(P, k) : tile_g(P, cr.ra, k); };
  __shared__ int s_ready;
  chain_wait(P, factor_flag(P, cr.ra, cr.k0), factor_flag(P, cr.jb, cr.k0));
  tile_load16(tileA(cr.k0), ra);
  tile_load16(tile_g(P, cr.jb, cr.k0), rb);
  for (int k = cr.k0;; ++k) {
    const bool more = k + 1 < cr.k1;
    __syncthreads();  // previous product done with XA / XB
    tile_store16(ra, XA);
    tile_store16(rb, XB);
    // The next contribution is prefetched only if its tiles are already
    // published (checked without waiting); otherwise this product runs
    // first and the wait follows it, so that the last product of a task is
    // not held behind the wait for it.
    if (threadIdx.x == 0) {
      int ready = 0;
      if (more &&
          __hip_atomic_load(factor_flag(P, cr.ra, k + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &&
          __hip_atomic_load(factor_flag(P, cr.jb, k + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ready = 1;
      }
      s_ready = ready;
    }
    __syncthreads();
    const bool pre = more && s_ready;
    if (pre) {
      tile_load16(tileA(k + 1), ra);
      tile_load16(tile_g(P, cr.jb, k + 1), rb);
    }
    mfma_abt(XA, XB, acc, sgn, lower_only);
    if (!more) break;
    if (!pre) {
      chain_wait(P, factor_flag(P, cr.ra, k + 1), factor_flag(P, cr.jb, k + 1));
      tile_load16(tileA(k + 1), ra);
      tile_load16(tile_g(P, cr.jb, k + 1), rb);
    }
  }
  __syncthreads();
}

template <int TAG>
__device__ __forceinline__ void helper_accumulate(const PobtafChain& P, const ContribRange& cr, double sgn,
                                  bool lower_only, double* XA, double* XB, f64x4 (&acc)[2][2]) {
  if (cr.k0 >= cr.k1) return;
  if (P.f32) {
    f32x4 a32[2][2];
    cacc32_zero(a32);
    helper_products(P, cr, sgn, lower_only, XA, XB, a32);
    acc32_add(a32, acc, XA, KP);
  } else {
    helper_products(P, cr, sgn, lower_only, XA, XB, acc);
  }
}

// out = acc * Linv(j)^T
template <int TAG>
__device__ __forceinline__ void helper_solve(const PobtafChain& P, int j, double* XA, double* XB,
                             f64x4 (&acc)[2][2]) {
  chain_wait(P, P.cflags + 2 * j, nullptr);
  TileRegs16 rl;
  tile_load16({P.linv + (int64_t)j * CT * CT, CT, CT, CT}, rl);
  cacc_to_lds(acc, XA, KP);  // A[m][k] = acc(m, k)
  tile_store16(rl, XB);      // B[n][k] = Linv(n, k)
  __syncthreads();
  cacc_zero(acc);
  if (P.f32) {
    f32x4 a32[2][2];
    cacc32_zero(a32);
    mfma_abt(XA, XB, a32, 1.0, false);
    acc32_add(a32, acc, XA, KP);
  } else {
    mfma_abt(XA, XB, acc, 1.0, false);
  }
}

// ------------------------------------------------------------- chain work
// LDS plan of the chain workgroup: S [64][65] (potrf, then L(g+1,g) staged
// K-contiguous with pitch 65 for the next step's update), Xs [64][65]
// (Linv(g)), Ws [64][17] (trtri scratch).
template <int TAG>
__device__ __forceinline__ void chain_run(const PobtafChain& P, double* S, double* Xs, double* Ws) {
  const int tid = threadIdx.x;
  double (*S2)[TT_LD] = reinterpret_cast<double (*)[TT_LD]>(S);
  double (*X2)[TT_LD] = reinterpret_cast<double (*)[TT_LD]>(Xs);
  double (*W2)[WS_LD] = reinterpret_cast<double (*)[WS_LD]>(Ws);
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  // Linv tiles keep zeros above their 16x16 block diagonal from here on
  // (full-tile inversions never write there; ragged ones clear it).
  for (int idx = tid; idx < CT * TT_LD; idx += CH_THREADS) Xs[idx] = 0.0;
  __syncthreads();
  // PD(g+1) is usually ready when the chain waits for PX(g): it is waited
  // for together with PX(g) and loaded into registers (pd_next) then, so
  // that the next step starts without the load latency.
  f64x4 pd_next[2][2];
  bool have_pd = false;
  __shared__ int s_early;
  // L(g+1,g) is stored at the end of step g but published during step g+1,
  // after its syrk products have read the tile from LDS: the publish
  // (every wave's store wait, a barrier, the release) then replaces the
  // barrier that follows the products instead of adding one, and the
  // release runs on wave 3, which has no work in the first panel of the
  // tile factorization that follows.
  int pending = -1;  // step whose L(g+1,g) stores are not yet published
  constexpr int PUBLISHER = 3 * 64;
  for (int g = 0; g < P.NT; ++g) {
    const CTile dt = tile_g(P, g, g);
    f64x4 acc[2][2];
    if (have_pd) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = pd_next[i][j];
    } else {
      if (g > 0) chain_wait(P, P.hflags + t_pd(P, g), nullptr);
      cacc_load(dt, acc);
    }
    have_pd = false;
    if (g > 0 && P.f32) {
      f32x4 a32[2][2];
      cacc32_zero(a32);
      if (wn <= wm) {
#pragma unroll 4
        for (int kk = 0; kk < CT; kk += 4) {
          const double a0 = -S[(wm + li) * TT_LD + kk + lk];
          const double a1 = -S[(wm + 16 + li) * TT_LD + kk + lk];
          const double b0 = S[(wn + li) * TT_LD + kk + lk];
          const double b1 = S[(wn + 16 + li) * TT_LD + kk + lk];
          a32[0][0] = mfma_acc(a0, b0, a32[0][0]);
          a32[0][1] = mfma_acc(a0, b1, a32[0][1]);
          a32[1][0] = mfma_acc(a1, b0, a32[1][0]);
          a32[1][1] = mfma_acc(a1, b1, a32[1][1]);
        }
      }
      if (pending >= 0) chain_publish(P.cflags + 2 * pending + 1, PUBLISHER);
      pending = -1;
      acc32_add(a32, acc, S, TT_LD);
    } else if (g > 0) {
      // acc -= L(g,g-1) L(g,g-1)^T, with L(g,g-1) in S as S[row][k]
      if (wn <= wm) {
#pragma unroll 4
        for (int kk = 0; kk < CT; kk += 4) {
          const double a0 = -S[(wm + li) * TT_LD + kk + lk];
          const double a1 = -S[(wm + 16 + li) * TT_LD + kk + lk];
          const double b0 = S[(wn + li) * TT_LD + kk + lk];
          const double b1 = S[(wn + 16 + li) * TT_LD + kk + lk];
          acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
          acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
          acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
          acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
        }
      }
      if (pending >= 0)
        chain_publish(P.cflags + 2 * pending + 1, PUBLISHER);  // includes the barrier
      else
        __syncthreads();
      pending = -1;
    }
    cacc_to_lds(acc, S, TT_LD);
    __syncthreads();
    const int m = dt.rows;
    const bool sub = g + 1 < band_end(P, g);
    const bool pre = sub && g + 1 < P.NT;
    const int* f_px = P.hflags + t_px(P, g);
    const int* f_pd = pre ? P.hflags + t_pd(P, g + 1) : f_px;
    // Full tiles: the inverse Linv(g) is formed by 16-row block rows while
    // the tile is factored.  Block row I needs only panels 0 .. I, so waves
    // 1-3 form block rows 0, 1, 2 while wave 0 factors panels 1, 2, 3
    // (X_II by forward substitution, X_IJ = -X_II sum_K L_IK X_KJ on MFMA);
    // block row 3 follows the factorization.  A wave that needs X_22 or X_33
    // before another wave's copy is visible forms its own copy in Ws.
    // PX(g) and PD(g+1) are normally published while the tile is factored.
    // Wave 1 checks their flags while wave 0 factors the last panel and, if
    // both are set, acquires; every thread then loads its parts of the two
    // tiles after its L(g,g) stores (which would otherwise queue behind the
    // loads), and the loads land during the rest of the inversion.
    {
      const int64_t row0 = (int64_t)(g / P.nt) * P.b + (int64_t)(g % P.nt) * CT;
      if (m == CT) {
        lds_potrf_t<CT>(S2, CT, P.info, row0, [&](int p) {
          if (p == 0 && wave == 3) wave_diag_inv16(S2, X2, 0);
          if (p == 16 && wave == 2) {
            wave_diag_inv16(S2, X2, 16);
            __builtin_amdgcn_wave_barrier();
            wave_trtri_offdiag(S2, X2, &X2[16][16], TT_LD, 1, 0);
          }
          if (p == 32) {
            if (wave == 1) {
              int ok = sub ? (__hip_atomic_load(f_px, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                              __hip_atomic_load(f_pd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                           : 0;
              ok = __builtin_amdgcn_readfirstlane(ok);
              if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
              if (lane == 0) s_early = ok;
            } else if (wave == 2) {
              wave_diag_inv16(S2, X2, 32);
              __builtin_amdgcn_wave_barrier();
              wave_trtri_offdiag(S2, X2, &X2[32][32], TT_LD, 2, 1);
            } else {
              wave_diag_inv16_to(S2, 32, Ws, WS_LD);
              __builtin_amdgcn_wave_barrier();
              wave_trtri_offdiag(S2, X2, Ws, WS_LD, 2, 0);
            }
          }
        });
      } else {
        if (tid == 0) s_early = 0;
        lds_potrf_t<0>(S2, m, P.info, row0);
      }
    }
    const int early = s_early;
    // L(g,g) to memory with its upper triangle zero.  Full tiles: each
    // thread reads its 8 row pairs from LDS first, then issues 16-byte
    // stores (rows rs + 8q, columns c, c + 1).
    const int rs = tid >> 5, c2 = (tid & 31) * 2;
    if (m == CT && (dt.ld & 1) == 0 && (reinterpret_cast<uintptr_t>(dt.p) & 15) == 0) {
      cdbl2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int rr = rs + 8 * q;
        v[q].x = (c2 <= rr) ? S2[rr][c2] : 0.0;
        v[q].y = (c2 + 1 <= rr) ? S2[rr][c2 + 1] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        *reinterpret_cast<cdbl2*>(dt.p + (int64_t)(rs + 8 * q) * dt.ld + c2) = v[q];
    } else {
      for (int idx = tid; idx < CT * CT; idx += CH_THREADS) {
        const int rr = idx >> 6, cc = idx & 63;
        if (rr < m && cc < m) dt.p[(int64_t)rr * dt.ld + cc] = (cc <= rr) ? S2[rr][cc] : 0.0;
      }
    }
    const CTile xt = tile_g(P, g + 1, g);
    TileRegs16 rx;
    if (early) {
      tile_load16(xt, rx);
      if (pre) cacc_load(tile_g(P, g + 1, g + 1), pd_next);
    }
    // Otherwise each wave reads both flags again now and checks them after
    // the Linv publish: when both were set it only acquires, without the
    // poll's round trip.
    int seen = 0;
    if (sub && !early)
      seen = __hip_atomic_load(f_px, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
             __hip_atomic_load(f_pd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (m == CT) {
      // block row 3: X_33 on waves 0-2 (wave 0's copy is the result), then
      // X_32, X_31, X_30
      if (wave < 3) {
        double* xi = (wave == 0) ? &X2[48][48] : Ws + 16 * wave * WS_LD;
        const int ldi = (wave == 0) ? TT_LD : WS_LD;
        wave_diag_inv16_to(S2, 48, xi, ldi);
        __builtin_amdgcn_wave_barrier();
        wave_trtri_offdiag(S2, X2, xi, ldi, 3, 2 - wave);
      }
      __syncthreads();
    } else {
      lds_trtri<false>(S2, X2, W2, m);
    }
    double* li_out = P.linv + (int64_t)g * CT * CT;
    {
      cdbl2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        v[q].x = X2[rs + 8 * q][c2];
        v[q].y = X2[rs + 8 * q][c2 + 1];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        *reinterpret_cast<cdbl2*>(li_out + (rs + 8 * q) * CT + c2) = v[q];
    }
    // Linv(g) is published before the sub-diagonal solve: the helpers'
    // solves of column g (TRSM(g+2,g) feeds PX(g+1), which the next step
    // waits for) start while the chain finishes this step.
    chain_publish(P.cflags + 2 * g);
    if (!sub) continue;
    // L(g+1,g) = PX(g) * Linv(g)^T.  Each thread loads only its own parts of
    // the two tiles, so the waves acquire independently; the barrier after
    // the staging below orders the LDS.
    if (!early) {
      if (__builtin_amdgcn_readfirstlane(seen))
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      else
        wave_wait(P, f_px, pre ? f_pd : nullptr);
      tile_load16(xt, rx);
      if (pre) cacc_load(tile_g(P, g + 1, g + 1), pd_next);
    }
    have_pd = pre;
    {
      // S[m][k] = X(m, k) (pitch 65 rows are 8-byte aligned only: scalar stores)
      const int rs = tid >> 5, c = (tid & 31) * 2;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        S[(rs + 8 * q) * TT_LD + c] = rx.v[q].x;
        S[(rs + 8 * q) * TT_LD + c + 1] = rx.v[q].y;
      }
    }
    __syncthreads();
    cacc_zero(acc);
    if (P.f32) {
      f32x4 a32[2][2];
      cacc32_zero(a32);
#pragma unroll 4
      for (int kk = 0; kk < CT; kk += 4) {
        const double a0 = S[(wm + li) * TT_LD + kk + lk];
        const double a1 = S[(wm + 16 + li) * TT_LD + kk + lk];
        const double b0 = Xs[(wn + li) * TT_LD + kk + lk];  // Linv(n, k)
        const double b1 = Xs[(wn + 16 + li) * TT_LD + kk + lk];
        a32[0][0] = mfma_acc(a0, b0, a32[0][0]);
        a32[0][1] = mfma_acc(a0, b1, a32[0][1]);
        a32[1][0] = mfma_acc(a1, b0, a32[1][0]);
        a32[1][1] = mfma_acc(a1, b1, a32[1][1]);
      }
      acc32_add(a32, acc, S, TT_LD);
    } else {
#pragma unroll 4
      for (int kk = 0; kk < CT; kk += 4) {
        const double a0 = S[(wm + li) * TT_LD + kk + lk];
        const double a1 = S[(wm + 16 + li) * TT_LD + kk + lk];
        const double b0 = Xs[(wn + li) * TT_LD + kk + lk];  // Linv(n, k)
        const double b1 = Xs[(wn + 16 + li) * TT_LD + kk + lk];
        acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
        acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
        acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
        acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
      }
    }
    __syncthreads();  // S no longer read as X
    cacc_to_lds(acc, S, TT_LD);  // keep L(g+1,g) as S[row][k]
    __syncthreads();
    lds_tile_store(S, TT_LD, xt);
    pending = g;
  }
  if (pending >= 0) chain_publish(P.cflags + 2 * pending + 1, PUBLISHER);
}


// Column block J of L(g+1,g) = PX(g) Linv(g)^T (rows 0..63, K < 16J + 16),
// with a compile-time trip count so that the LDS operand loads are
// scheduled ahead of the products.  ACC f64x4: FP64 MFMA; f32x4: FP32.
template <int J, typename ACC>
__device__ __forceinline__ void chain_colblock(const double* Ps, const double (*X2)[TT_LD],
                                               ACC (&cb)[4]) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) cb[rb] = ACC{0, 0, 0, 0};
#pragma unroll
  for (int kk = 0; kk < 16 * J + 16; kk += 4) {
    const double bv = X2[16 * J + li][kk + lk];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) cb[rb] = mfma_acc(Ps[(16 * rb + li) * TT_LD + kk + lk], bv, cb[rb]);
  }
}

// ------------------------------------------------------- fused chain step
// Chain schedule for matrices whose diagonal tiles are all full (b % 64 ==
// 0).  Step g starts with D(g,g) fully updated in S and:
//   1. factors it (wave 0 factors the 16-column panels; waves 1-3 update the
//      trailing columns and form block rows 0-2 of Linv(g) beside them);
//   2. with PX(g) staged in Ps, waves 1-3 form column blocks 0-2 of
//      L(g+1,g) = PX(g) Linv(g)^T (column block J needs only block row J of
//      Linv, already formed), while wave 0 forms X_33 = L(g,g)[48:64,48:64]^{-1};
//   3. forms column block 3 by substitution (every wave 16 rows):
//        L(g+1,g)[:, 48:64] = (PX(g)[:, 48:64] - L(g+1,g)[:, 0:48] L(g,g)[48:64, 0:48]^T) X_33^T,
//      and waves 1-3 then form X_30, X_31, X_32, the rest of Linv's block row 3;
//   4. D(g+1,g+1) = PD(g+1) - L(g+1,g) L(g+1,g)^T on waves 0, 2, 3 (the lower
//      quadrants) into the other D buffer.
// Wave 1 stores Linv(g) after step 3 and releases it during step 4, where
// it has no quadrant; L(g+1,g) is released at the start of step g+1, whose
// first panel only wave 0 works on.  In the mixed mode the products
// (column blocks, the column-block-3 update and its X_33 product, the syrk)
// round their operands to FP32 and accumulate in FP32; the differences and
// the PD(g+1) term stay FP64.
// LDS: two D buffers [64][65] (D(g,g) -> L(g,g), and D(g+1,g+1)), Xs
// [64][65] (Linv(g)), Ws [64][17] (scratch), Ps [64][65] (PX(g) -> L(g+1,g),
// row-major = K-contiguous).
template <int TAG>
__device__ __forceinline__ void chain_run_fused(const PobtafChain& P, double* Sa, double* Sb,
                                                double* Xs, double* Ws, double* Ps) {
  const int tid = threadIdx.x;
  double (*X2)[TT_LD] = reinterpret_cast<double (*)[TT_LD]>(Xs);
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int rs = tid >> 5, c2 = (tid & 31) * 2;
  constexpr int PUB = 64;  // wave 1, lane 0: runs the releases
  for (int idx = tid; idx < CT * TT_LD; idx += CH_THREADS) Xs[idx] = 0.0;
  {
    f64x4 acc[2][2];
    cacc_load(tile_g(P, 0, 0), acc);  // D(0,0) has no updates
    cacc_to_lds(acc, Sa, TT_LD);
  }
  __syncthreads();
  f64x4 pd_next[2][2];
  __shared__ int s_early;
  int pending = -1;  // step whose Linv and L(g+1,g) stores wait for their release
  double* S = Sa;
  double* Sn = Sb;
  for (int g = 0; g < P.NT; ++g) {
    double (*S2)[TT_LD] = reinterpret_cast<double (*)[TT_LD]>(S);
    const CTile dt = tile_g(P, g, g);
    const bool sub = g + 1 < band_end(P, g);
    const int* f_px = P.hflags + t_px(P, g);
    const int* f_pd = sub ? P.hflags + t_pd(P, g + 1) : f_px;
    if (pending >= 0 && tid == PUB) {
      // every wave waited for its L(g,g-1) stores before the barrier that
      // ended the previous step
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      wave_stores_done();
      __hip_atomic_store(P.cflags + 2 * pending + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    pending = -1;
    chain_trace(P, g, 0);
    // 1. factorization with Linv block rows 0-2 beside it; PX(g) and
    //    PD(g+1) flags checked by wave 1 during the last panel
    const int64_t row0 = (int64_t)(g / P.nt) * P.b + (int64_t)(g % P.nt) * CT;
    lds_potrf_t<CT>(S2, CT, P.info, row0, [&](int p) {
      if (p == 0 && wave == 3) wave_diag_inv16(S2, X2, 0);
      if (p == 16 && wave == 2) {
        wave_diag_inv16(S2, X2, 16);
        __builtin_amdgcn_wave_barrier();
        wave_trtri_offdiag(S2, X2, &X2[16][16], TT_LD, 1, 0);
      }
      if (p == 32) {
        if (wave == 1) {
          int ok = sub ? (__hip_atomic_load(f_px, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) &
                          __hip_atomic_load(f_pd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                       : 0;
          ok = __builtin_amdgcn_readfirstlane(ok);
          if (ok) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          if (lane == 0) s_early = ok;
        } else if (wave == 2) {
          wave_diag_inv16(S2, X2, 32);
          __builtin_amdgcn_wave_barrier();
          wave_trtri_offdiag(S2, X2, &X2[32][32], TT_LD, 2, 1);
        } else {
          wave_diag_inv16_to(S2, 32, Ws, WS_LD);
          __builtin_amdgcn_wave_barrier();
          wave_trtri_offdiag(S2, X2, Ws, WS_LD, 2, 0);
        }
      }
    });
    const int early = s_early;
    chain_trace(P, g, 1);
    // L(g,g) to memory, upper triangle zero (16-byte stores)
    {
      cdbl2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int rr = rs + 8 * q;
        v[q].x = (c2 <= rr) ? S2[rr][c2] : 0.0;
        v[q].y = (c2 + 1 <= rr) ? S2[rr][c2 + 1] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        *reinterpret_cast<cdbl2*>(dt.p + (int64_t)(rs + 8 * q) * dt.ld + c2) = v[q];
    }
    double* li_out = P.linv + (int64_t)g * CT * CT;
    if (!sub) {
      // last diagonal tile of the matrix: Linv block row 3 and its release
      if (wave < 3) {
        double* xi = (wave == 0) ? &X2[48][48] : Ws + 16 * wave * WS_LD;
        const int ldi = (wave == 0) ? TT_LD : WS_LD;
        wave_diag_inv16_to(S2, 48, xi, ldi);
        __builtin_amdgcn_wave_barrier();
        wave_trtri_offdiag(S2, X2, xi, ldi, 3, 2 - wave);
      }
      __syncthreads();
      cdbl2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        v[q].x = X2[rs + 8 * q][c2];
        v[q].y = X2[rs + 8 * q][c2 + 1];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) *reinterpret_cast<cdbl2*>(li_out + (rs + 8 * q) * CT + c2) = v[q];
      chain_publish(P.cflags + 2 * g);
      continue;
    }
    // PX(g) and PD(g+1): loaded now if both were published during the
    // factorization, else each wave waits for them itself
    const CTile xt = tile_g(P, g + 1, g);
    TileRegs16 rx;
    if (!early) wave_wait(P, f_px, f_pd);
    tile_load16(xt, rx);
    cacc_load(tile_g(P, g + 1, g + 1), pd_next);
    // Ps[m][k] = PX(g)(m, k) (pitch 65: 8-byte stores)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      Ps[(rs + 8 * q) * TT_LD + c2] = rx.v[q].x;
      Ps[(rs + 8 * q) * TT_LD + c2 + 1] = rx.v[q].y;
    }
    __syncthreads();
    chain_trace(P, g, 2);
    // 2. wave 0: X_33 = L(g,g)[48:64, 48:64]^{-1};
    //    wave J+1: column block J of L(g+1,g) = sum over k < 16J + 16 of
    //    PX(:, k) Linv(16J + n, k)
    f64x4 cb[4];
    f32x4 cb32[4];
    if (wave == 0) {
      wave_diag_inv16(S2, X2, 48);
    } else if (!P.f32) {
      if (wave == 1) chain_colblock<0>(Ps, X2, cb);
      else if (wave == 2) chain_colblock<1>(Ps, X2, cb);
      else chain_colblock<2>(Ps, X2, cb);
    } else {
      if (wave == 1) chain_colblock<0>(Ps, X2, cb32);
      else if (wave == 2) chain_colblock<1>(Ps, X2, cb32);
      else chain_colblock<2>(Ps, X2, cb32);
    }
    __syncthreads();  // every read of PX(g)'s column blocks 0-2 is done; X_33 is visible
    if (wave >= 1) {
      const int J = wave - 1;
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (!P.f32)
            Ps[(16 * rb + acc_row_f64(lane, r)) * TT_LD + 16 * J + acc_col(lane)] = cb[rb][r];
          else
            Ps[(16 * rb + acc_row_f32(lane, r)) * TT_LD + 16 * J + acc_col(lane)] = (double)cb32[rb][r];
        }
    }
    __syncthreads();  // column blocks 0-2 of L(g+1,g) in Ps
    chain_trace(P, g, 3);
    // 3. column block 3, wave w: rows 16w .. 16w + 15.  Ws receives the
    //    product L(g+1,g)[:, 0:48] L(g,g)[48:64, 0:48]^T; the operand of the
    //    X_33 product is PX(g)[:, 48:64] minus it, formed in FP64.
    {
      const int r0 = 16 * wave;
      if (!P.f32) {
        f64x4 v = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 48; kk += 4)
          v = mfma_f64(Ps[(r0 + li) * TT_LD + kk + lk], S[(48 + li) * TT_LD + kk + lk], v);
#pragma unroll
        for (int r = 0; r < 4; ++r) Ws[(r0 + acc_row_f64(lane, r)) * WS_LD + acc_col(lane)] = v[r];
      } else {
        f32x4 v = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 48; kk += 4)
          v = mfma_acc(Ps[(r0 + li) * TT_LD + kk + lk], S[(48 + li) * TT_LD + kk + lk], v);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Ws[(r0 + acc_row_f32(lane, r)) * WS_LD + acc_col(lane)] = (double)v[r];
      }
      __builtin_amdgcn_wave_barrier();
      if (!P.f32) {
        f64x4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
          const double av = Ps[(r0 + li) * TT_LD + 48 + kk + lk] - Ws[(r0 + li) * WS_LD + kk + lk];
          y = mfma_f64(av, X2[48 + li][48 + kk + lk], y);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) Ps[(r0 + acc_row_f64(lane, r)) * TT_LD + 48 + acc_col(lane)] = y[r];
      } else {
        f32x4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 16; kk += 4) {
          const double av = Ps[(r0 + li) * TT_LD + 48 + kk + lk] - Ws[(r0 + li) * WS_LD + kk + lk];
          y = mfma_acc(av, X2[48 + li][48 + kk + lk], y);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
          Ps[(r0 + acc_row_f32(lane, r)) * TT_LD + 48 + acc_col(lane)] = (double)y[r];
      }
    }
    // the rest of Linv's block row 3: wave J+1 forms X_3J
    if (wave >= 1) wave_trtri_offdiag(S2, X2, &X2[48][48], TT_LD, 3, wave - 1);
    __syncthreads();  // L(g+1,g) complete in Ps; Linv(g) complete
    chain_trace(P, g, 4);
    if (wave == 1) {
      // Linv(g) to memory by wave 1 alone (released by it in step 4), two
      // whole rows per store instruction
      const int hr = lane >> 5, cc = (lane & 31) * 2;
#pragma unroll 8
      for (int k = 0; k < 32; ++k) {
        const int row = 2 * k + hr;
        cdbl2 v;
        v.x = X2[row][cc];
        v.y = X2[row][cc + 1];
        *reinterpret_cast<cdbl2*>(li_out + row * CT + cc) = v;
      }
    }
    lds_tile_store(Ps, TT_LD, xt);
    pending = g;
    // 4. D(g+1,g+1) = PD(g+1) - L(g+1,g) L(g+1,g)^T on the lower quadrants
    //    (waves 0, 2, 3) into Sn; wave 1, which has no lower quadrant, waits
    //    for its Linv(g) stores and releases them
    f32x4 s32[2][2];
    cacc32_zero(s32);
    if (wn <= wm) {
      if (!P.f32) {
#pragma unroll
        for (int kk = 0; kk < CT; kk += 4) {
          const double a0 = -Ps[(wm + li) * TT_LD + kk + lk];
          const double a1 = -Ps[(wm + 16 + li) * TT_LD + kk + lk];
          const double b0 = Ps[(wn + li) * TT_LD + kk + lk];
          const double b1 = Ps[(wn + 16 + li) * TT_LD + kk + lk];
          pd_next[0][0] = mfma_f64(a0, b0, pd_next[0][0]);
          pd_next[0][1] = mfma_f64(a0, b1, pd_next[0][1]);
          pd_next[1][0] = mfma_f64(a1, b0, pd_next[1][0]);
          pd_next[1][1] = mfma_f64(a1, b1, pd_next[1][1]);
        }
      } else {
#pragma unroll
        for (int kk = 0; kk < CT; kk += 4) {
          const double a0 = -Ps[(wm + li) * TT_LD + kk + lk];
          const double a1 = -Ps[(wm + 16 + li) * TT_LD + kk + lk];
          const double b0 = Ps[(wn + li) * TT_LD + kk + lk];
          const double b1 = Ps[(wn + 16 + li) * TT_LD + kk + lk];
          s32[0][0] = mfma_acc(a0, b0, s32[0][0]);
          s32[0][1] = mfma_acc(a0, b1, s32[0][1]);
          s32[1][0] = mfma_acc(a1, b0, s32[1][0]);
          s32[1][1] = mfma_acc(a1, b1, s32[1][1]);
        }
      }
    } else {
      wave_stores_done();
      if (lane == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        wave_stores_done();
        __hip_atomic_store(P.cflags + 2 * g, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    // FP32 products join the FP64 PD(g+1) through Sn (free until below)
    if (P.f32) acc32_add(s32, pd_next, Sn, TT_LD);
    cacc_to_lds(pd_next, Sn, TT_LD);
    wave_stores_done();  // Linv(g), L(g,g), L(g+1,g): released at the start of step g+1
    __syncthreads();
    chain_trace(P, g, 5);
    double* t = S;
    S = Sn;
    Sn = t;
  }
  if (pending >= 0 && tid == PUB) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    wave_stores_done();
    __hip_atomic_store(P.cflags + 2 * pending + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------------------------------------------------------- kernel
// One helper task (already claimed) of matrix P.
template <int TAG>
__device__ __forceinline__ void chain_helper_task(const PobtafChain& P, int task, double* lds) {
  const int tid = threadIdx.x;
  double* XA = lds;
  double* XB = lds + CT * KP;
  int r, j;
  const int kind = chain_decode(P, task, r, j);
  if (kind == HK_NOOP) return;  // nobody waits on an empty slot

  f64x4 acc[2][2];
  if (kind == HK_TIP) {
    cacc_load({P.T, P.a, P.a, P.a}, acc);
    for (int i = 0; i < P.n; ++i) {
      chain_wait(P, P.hflags + t_tippc(P, i), nullptr);
      f64x4 pi[2][2];
      cacc_load({P.tipp + (int64_t)i * P.a * P.a, P.a, P.a, P.a}, pi);
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] -= pi[x][y];
    }
    double (*S2)[TT_LD] = reinterpret_cast<double (*)[TT_LD]>(lds);
    cacc_to_lds(acc, lds, TT_LD);
    __syncthreads();
    lds_potrf(S2, P.a, P.info, (int64_t)P.n * P.b);
    for (int idx = tid; idx < P.a * P.a; idx += CH_THREADS) {
      const int rr = idx / P.a, cc = idx % P.a;
      P.T[idx] = (cc <= rr) ? S2[rr][cc] : 0.0;
    }
    if (P.tip_inv) {
      double (*X2)[TT_LD] = reinterpret_cast<double (*)[TT_LD]>(lds + CT * TT_LD);
      double (*W2)[WS_LD] = reinterpret_cast<double (*)[WS_LD]>(lds + 2 * CT * TT_LD);
      __syncthreads();
      lds_trtri(S2, X2, W2, P.a);
      for (int idx = tid; idx < CT * CT; idx += CH_THREADS) {
        const int rr = idx / CT, cc = idx % CT;
        P.tip_inv[idx] = (rr < P.a && cc <= rr) ? X2[rr][cc] : 0.0;
      }
    }
    chain_publish(P.hflags + task);
    return;
  }
  if (kind == HK_TIPP) {
    const int i = j;
    cacc_zero(acc);
    for (int c = 0; c < P.nt; ++c) {
      const int k = i * P.nt + c;
      chain_wait(P, P.hflags + t_arr(P, k), nullptr);
      TileRegs16 x;
      tile_load16(tile_arrow(P, k), x);
      __syncthreads();
      tile_store16(x, XA);
      __syncthreads();
      mfma_abt(XA, XA, acc, 1.0, true);
    }
    cacc_store({P.tipp + (int64_t)i * P.a * P.a, P.a, P.a, P.a}, acc);  // FP64 (tip path)
    chain_publish(P.hflags + task);
    return;
  }
  CTile out;
  if (kind == HK_TRSM)
    out = tile_g(P, r, j);
  else if (kind == HK_ARR)
    out = tile_arrow(P, j);
  else if (kind == HK_PX)
    out = tile_g(P, j + 1, j);
  else
    out = tile_g(P, j, j);  // PD
  cacc_load(out, acc);
  const ContribRange cr = chain_contribs(P, kind, r, j);
  helper_accumulate<TAG>(P, cr, -1.0, kind == HK_PD, XA, XB, acc);
  if (kind == HK_TRSM || kind == HK_ARR) helper_solve<TAG>(P, j, XA, XB, acc);
  __syncthreads();  // the last product is done with XA
  cacc_to_lds(acc, XA, KP);
  __syncthreads();
  lds_tile_store(XA, KP, out);
  chain_publish(P.hflags + task);
  if (kind == HK_PX) chain_trace(P, j, 6);
  if (kind == HK_TRSM && r == j + 2) chain_trace(P, j, 7);
  // The strict-upper mirror of a lower D-block tile is zero in the factor.
  // No task reads it, so it is cleared after the publish, off the path of
  // the tasks waiting for this one.
  if (kind == HK_TRSM && r / P.nt == j / P.nt) tile_zero(tile_mirror(P, r, j));
  if (kind == HK_PX && (j + 1) / P.nt == j / P.nt) tile_zero(tile_mirror(P, j + 1, j));
}

// LDS of the single-matrix kernel: the helper plan (XA, XB) or the chain's
// (S, Xs, Ws, and Ps for the fused schedule), whichever is larger.
constexpr int CH_LDS_HELPER = 2 * CT * KP + CT * WS_LD;
constexpr int CH_LDS_FUSED = 4 * CT * TT_LD + CT * WS_LD;
constexpr int CH_LDS = CH_LDS_HELPER > CH_LDS_FUSED ? CH_LDS_HELPER : CH_LDS_FUSED;

__global__ void __launch_bounds__(CH_THREADS) pobtaf_chain_kernel(PobtafChain P) {
  __shared__ __attribute__((aligned(16))) double lds[CH_LDS];
  __shared__ int s_task;
  const int tid = threadIdx.x;
  if (tid == 0) s_task = atomicAdd(P.claim, 1);
  __syncthreads();
  const bool is_chain = (s_task == 0);
  __syncthreads();
  if (is_chain) {
    // the chain runs latency-bound code: keep helpers off its CU
    if (tid == 0)
      __hip_atomic_store(P.chain_cu, 1 + hw_cu_key(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (P.fused)
      chain_run_fused<0>(P, lds, lds + CT * TT_LD, lds + 2 * CT * TT_LD, lds + 3 * CT * TT_LD,
                         lds + 3 * CT * TT_LD + CT * WS_LD);
    else
      chain_run<0>(P, lds, lds + CT * TT_LD, lds + 2 * CT * TT_LD);
    return;
  }
  const int my_cu = 1 + hw_cu_key();
  for (;;) {
    if (tid == 0) {
      const int ccu = __hip_atomic_load(P.chain_cu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_task = (ccu == my_cu) ? P.ntasks : atomicAdd(P.counter, 1);
    }
    __syncthreads();
    const int task = s_task;
    __syncthreads();
    if (task >= P.ntasks) return;
    chain_helper_task<0>(P, task, lds);
  }
}

// Many independent matrices in one persistent launch (batched plans).  The
// first nmat claims of *gclaim become the chains of matrices 0..nmat-1;
// every other workgroup serves the matrices' helper queues round-robin,
// taking each task from that matrix's own in-order counter, and leaves when
// a full round finds every queue drained.  cu_busy (4096 ints, zero on
// entry) marks the CUs that run a chain; helpers there leave at once.
__global__ void __launch_bounds__(CH_THREADS) pobtaf_chain_multi_kernel(const PobtafChain* Ps,
                                                                        int nmat, int* gclaim,
                                                                        int* cu_busy) {
  __shared__ __attribute__((aligned(16))) double lds[2 * CT * KP + CT * WS_LD];
  __shared__ int s_task;
  const int tid = threadIdx.x;
  if (tid == 0) s_task = atomicAdd(gclaim, 1);
  __syncthreads();
  const int first = s_task;
  __syncthreads();
  const int my_key = hw_cu_key() & 4095;
  if (first < nmat) {
    if (tid == 0) __hip_atomic_store(cu_busy + my_key, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const PobtafChain& P = Ps[first];
    chain_run<1>(P, lds, lds + CT * TT_LD, lds + 2 * CT * TT_LD);
    return;
  }
  int m = (int)(blockIdx.x % (unsigned)nmat);
  int drained = 0;  // consecutive matrices found without tasks
  while (drained < nmat) {
    const PobtafChain& P = Ps[m];
    if (tid == 0) {
      const int busy = __hip_atomic_load(cu_busy + my_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_task = busy ? -1 : atomicAdd(P.counter, 1);
    }
    __syncthreads();
    const int task = s_task;
    __syncthreads();
    if (task < 0) return;
    if (task >= P.ntasks) {
      ++drained;
    } else {
      drained = 0;
      chain_helper_task<1>(P, task, lds);
    }
    m = (m + 1 == nmat) ? 0 : m + 1;
  }
}

}  // namespace rocbta
