// Host orchestration and C ABI of the gfx950 BTA library (librocbta.so).
//
// Algorithms (reference recurrences; the kernels evaluate them tile by tile):
//   pobtaf  : src/serinv/algs/pobtaf.py:97-183 (block Cholesky)
//   pobtasi : src/serinv/algs/pobtasi.py:92-172 (selected inversion)
//   pobtas  : src/serinv/algs/pobtas.py:65-161 (triangular solves)
//   pobtf / pobtsi / pobts: the same with a == 0 (pobtf.py:83-122,
//   pobtsi.py:78-125, pobts.py:57-118).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "../../include/rocbta.h"
#include "gemm_f64.hpp"
#include "grad_f64.hpp"
#include "pobtaf_chain.hpp"
#include "solve_f64.hpp"
#include "solve_sweep.hpp"
#include "solve_block.hpp"
#include "solve_wide.hpp"
#include "resid_f64.hpp"
#include "tile_f64.hpp"
#include "getri_f64.hpp"

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <mutex>

using namespace rocbta;

namespace {

#define RB_CHECK(expr)                                   \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) {                              \
      fprintf(stderr, "rocbta: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(e_), \
              __FILE__, __LINE__);                       \
      return ROCBTA_ERR_HIP;                             \
    }                                                    \
  } while (0)

inline int cdiv(int64_t x, int64_t y) { return (int)((x + y - 1) / y); }
inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// ---------------------------------------------------------------- GEMM group
struct GemmList {
  std::vector<GemmProblem> probs;
  GemmList() { probs.reserve(8); }

  GemmProblem& problems_back() { return probs.back(); }
  GemmProblem& add(double* C, int64_t ldc, int M, int N, double beta, int uplo = TRI_NONE,
                   int batch1 = 1, int64_t sC1 = 0, int batch2 = 1, int64_t sC2 = 0) {
    GemmProblem p{};
    p.C = C;
    p.ldc = ldc;
    p.sC1 = sC1;
    p.sC2 = sC2;
    p.M = M;
    p.N = N;
    p.batch1 = batch1;
    p.batch2 = batch2;
    p.beta = beta;
    p.nterms = 0;
    p.uplo = uplo;
    p.ksplit = 1;
    probs.push_back(p);
    return probs.back();
  }
};

GemmTerm term(const double* A, int64_t lda, bool tA, const double* B, int64_t ldb, bool tB, int K,
              double alpha, int triA = TRI_NONE, int triB = TRI_NONE, int64_t sA1 = 0,
              int64_t sB1 = 0, int64_t sA2 = 0, int64_t sB2 = 0) {
  GemmTerm t{};
  t.A = A;
  t.B = B;
  t.lda = lda;
  t.ldb = ldb;
  t.sA1 = sA1;
  t.sA2 = sA2;
  t.sB1 = sB1;
  t.sB2 = sB2;
  t.K = K;
  t.transA = tA;
  t.transB = tB;
  t.triA = (uint8_t)triA;
  t.triB = (uint8_t)triB;
  t.alpha = alpha;
  return t;
}

// A problem holds at most two terms; a third marks it invalid (nterms = -1)
// and run_gemms rejects the list instead of launching it.
void push_term(GemmProblem& p, const GemmTerm& t) {
  if (t.K <= 0) return;
  if (p.nterms < 0 || p.nterms >= 2) {
    p.nterms = -1;
    return;
  }
  p.t[p.nterms++] = t;
}

// ----------------------------------------------------------- op recording
// A batched plan records the launches of many independent per-matrix op
// sequences instead of issuing them: op k of every matrix goes to phase k,
// and each phase later runs as (at most) one fill, one tile and one GEMM
// launch over device-resident descriptor lists.
struct FillOp {
  double* p;
  int64_t count;
  double v;
};
struct Phase {
  std::vector<FillOp> fills;
  std::vector<TileBatch> tiles;
  std::vector<GemmProblem> gemms;
};
struct Recorder {
  std::vector<Phase> phases;
  size_t cur = 0;
  Phase& next() {
    if (cur >= phases.size()) phases.emplace_back();
    return phases[cur++];
  }
};
thread_local Recorder* g_rec = nullptr;
// Set for the duration of a mixed-precision entry point: every GEMM problem
// and the persistent factorization run their products in FP32.
thread_local bool g_mixed = false;
struct MixedScope {
  bool prev;
  explicit MixedScope(bool on) : prev(g_mixed) { g_mixed = on; }
  ~MixedScope() { g_mixed = prev; }
};

int device_cus();  // CUs of the current device (below)

// Output tile size for a launch list.  64 by default: measured on the
// selected inversion at n=256, b=1024 the 128-tile kernel (one workgroup
// per CU, split-K 4) took 107.6 ms against 62.1 ms for the 64-tile kernel
// (two workgroups per CU).  ROCBTA_GEMM_TILE=128 selects it for lists whose
// largest problem is large in both output dimensions and in K.
bool gemm_tile128_enabled() {
  const char* e = getenv("ROCBTA_GEMM_TILE");
  return e && strcmp(e, "128") == 0;
}
int gemm_ts_for(int M, int N, int K) {
  return (gemm_tile128_enabled() && M >= 256 && N >= 256 && K >= 256) ? 128 : 64;
}
int gemm_ts(const GemmList& list) {
  int ts = 64;
  for (const GemmProblem& p : list.probs) {
    int kmax = 0;
    for (int t = 0; t < p.nterms; ++t) kmax = std::max(kmax, p.t[t].K);
    if (gemm_ts_for(p.M, p.N, kmax) == 128) ts = 128;
  }
  return ts;
}

int run_gemms(GemmList& list, hipStream_t s) {
  for (const GemmProblem& p : list.probs)
    if (p.nterms < 0) return ROCBTA_ERR_ARG;
  if (g_rec) {
    Phase& ph = g_rec->next();
    for (GemmProblem p : list.probs) {
      if (p.M <= 0 || p.N <= 0 || p.batch1 <= 0 || p.batch2 <= 0) continue;
      if (p.nterms == 0 && p.beta == 1.0) continue;
      p.tiles_m = cdiv(p.M, GT_M);
      p.tiles_n = cdiv(p.N, GT_N);
      p.ksplit = 1;
      p.partial = nullptr;
      p.counters = nullptr;
      p.f32 = g_mixed ? 1 : 0;
      ph.gemms.push_back(p);
    }
    return ROCBTA_OK;
  }
  const int ts = gemm_ts(list);
  // 512-thread 64-tile kernel for launches of at most one tile per CU (two
  // waves per SIMD instead of one); ROCBTA_GEMM_H2=0 / 1 forces it off / on.
  static const int h2_mode = [] {
    const char* e = getenv("ROCBTA_GEMM_H2");
    return e ? (strcmp(e, "1") == 0 ? 1 : 0) : -1;
  }();
  const int cus = device_cus();
  GemmGroup g{};
  auto flush = [&]() -> int {
    if (g.nprob == 0) return ROCBTA_OK;
    if (g.total_tiles > 0) {
      if (ts == 128)
        gemm_group_kernel_t128<<<g.total_tiles, GT_THREADS, 0, s>>>(g);
      else if (h2_mode == 1 || (h2_mode < 0 && g.total_tiles <= cus))
        gemm_group_kernel_h2<<<g.total_tiles, 2 * GT_THREADS, 0, s>>>(g);
      else
        gemm_group_kernel<<<g.total_tiles, GT_THREADS, 0, s>>>(g);
      RB_CHECK(hipGetLastError());
    }
    g = GemmGroup{};
    return ROCBTA_OK;
  };
  for (GemmProblem p : list.probs) {
    if (p.M <= 0 || p.N <= 0 || p.batch1 <= 0 || p.batch2 <= 0) continue;
    if (p.nterms == 0 && p.beta == 1.0) continue;
    p.tiles_m = cdiv(p.M, ts);
    p.tiles_n = cdiv(p.N, ts);
    p.f32 = g_mixed ? 1 : 0;
    if (p.ksplit < 1 || !p.partial || !p.counters) p.ksplit = 1;
    p.lower_packed = (p.uplo == TRI_LOWER && p.tiles_m == p.tiles_n) ? 1 : 0;
    const int64_t per_batch =
        p.lower_packed ? (int64_t)p.tiles_m * (p.tiles_m + 1) / 2 : (int64_t)p.tiles_m * p.tiles_n;
    const int64_t tiles = per_batch * p.batch1 * p.batch2 * p.ksplit;
    if (g.nprob == ROCBTA_MAX_PROBLEMS || (int64_t)g.total_tiles + tiles > (1ll << 30)) {
      int rc = flush();
      if (rc) return rc;
    }
    p.tile_begin = g.total_tiles;
    g.p[g.nprob++] = p;
    g.total_tiles += (int)tiles;
  }
  return flush();
}

__global__ void fill_i32_kernel(int* p, int64_t count, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

int fill(double* p, int64_t count, double v, hipStream_t s) {
  if (g_rec) {
    Phase& ph = g_rec->next();
    if (count > 0) ph.fills.push_back({p, count, v});
    return ROCBTA_OK;
  }
  if (count <= 0) return ROCBTA_OK;
  int blocks = (int)std::min<int64_t>(4096, (count + 255) / 256);
  fill_f64_kernel<<<blocks, 256, 0, s>>>(p, count, v);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

// ------------------------------------------------------------ tile kernels
int tile_factor(int mode, double* A, int64_t lda, int64_t sA1, int64_t sA2, double* X,
                int64_t ldx, int64_t sX1, int64_t sX2, int m, int batch1, int batch2, int* info,
                int64_t info_row0, int64_t info_s1, int64_t info_s2, int zcols,
                hipStream_t s) {
  if (m <= 0 || batch1 <= 0 || batch2 <= 0) {
    if (g_rec) g_rec->next();
    return ROCBTA_OK;
  }
  TileBatch tb{};
  tb.A = A;
  tb.lda = lda;
  tb.sA1 = sA1;
  tb.sA2 = sA2;
  tb.X = X;
  tb.ldx = ldx;
  tb.sX1 = sX1;
  tb.sX2 = sX2;
  tb.m = m;
  tb.batch1 = batch1;
  tb.batch2 = batch2;
  tb.info = info;
  tb.info_row0 = info_row0;
  tb.info_s1 = info_s1;
  tb.info_s2 = info_s2;
  tb.zcols = zcols;
  tb.mode = mode;
  if (g_rec) {
    g_rec->next().tiles.push_back(tb);
    return ROCBTA_OK;
  }
  const int grid = batch1 * batch2;
  if (mode == 0)
    tile_factor_kernel<0><<<grid, TT_THREADS, 0, s>>>(tb);
  else
    tile_factor_kernel<1><<<grid, TT_THREADS, 0, s>>>(tb);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

// ------------------------------------------------ blocked diagonal factor
struct Panel {
  double* p;
  int rows;
  int64_t ld;
};

// Cholesky of the m x m diagonal block Dm (row stride ld), with the
// triangular solves of the panels below it (each rows x m, row stride ld_p)
// done tile column by tile column: right-looking, 64-wide tiles.
// tinv: 64 x 64 device scratch for the inverse of the current diagonal tile.
int factor_block_column(double* Dm, int64_t ld, int m, const std::vector<Panel>& panels,
                        int64_t row0, int* info, double* tinv, hipStream_t s) {
  for (int k0 = 0; k0 < m; k0 += TT) {
    const int tk = std::min(TT, m - k0);
    const int rem = m - k0 - tk;
    double* dkk = Dm + (int64_t)k0 * ld + k0;
    int rc = tile_factor(0, dkk, ld, 0, 0, tinv, TT, 0, 0, tk, 1, 1, info, row0 + k0, 0, 0, rem,
                         s);
    if (rc) return rc;
    // X <- X * inv(L_kk)^T for every row below the diagonal tile (in place;
    // N = tk <= 64, so each workgroup owns whole rows).
    GemmList trsm;
    if (rem > 0) {
      double* c = Dm + (int64_t)(k0 + tk) * ld + k0;
      push_term(trsm.add(c, ld, rem, tk, 0.0), term(c, ld, false, tinv, TT, true, tk, 1.0,
                                                    TRI_NONE, TRI_UPPER));
    }
    for (const Panel& pn : panels) {
      if (pn.rows <= 0) continue;
      double* c = pn.p + k0;
      push_term(trsm.add(c, pn.ld, pn.rows, tk, 0.0),
                term(c, pn.ld, false, tinv, TT, true, tk, 1.0, TRI_NONE, TRI_UPPER));
    }
    rc = run_gemms(trsm, s);
    if (rc) return rc;
    if (rem <= 0) continue;
    // right-looking update of the trailing columns
    GemmList upd;
    const double* pd = Dm + (int64_t)(k0 + tk) * ld + k0;  // rem x tk
    push_term(upd.add(Dm + (int64_t)(k0 + tk) * (ld + 1), ld, rem, rem, 1.0, TRI_LOWER),
              term(pd, ld, false, pd, ld, true, tk, -1.0));
    for (const Panel& pn : panels) {
      if (pn.rows <= 0) continue;
      push_term(upd.add(pn.p + k0 + tk, pn.ld, pn.rows, rem, 1.0),
                term(pn.p + k0, pn.ld, false, pd, ld, true, tk, -1.0));
    }
    rc = run_gemms(upd, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

// Inverse of `batch` lower-triangular m x m matrices (source stride sL,
// row stride lds) into X (row stride ldx, stride sX), upper part zeroed.
// W: scratch of batch * ceil(m/64) * 64 * 64 doubles.
// zero_upper = false (halving path, 64-tile products only): the strictly
// upper 64-tiles of X are left unwritten.  The halving recursion reads X's
// blocks only as lower triangular operands, whose products clip their K
// range at the output tile and so skip those tiles, and the diagonal tiles
// come zero above their diagonal from the tile kernel; callers that read X
// only on and below the diagonal tiles save the n b^2 fill.
int block_trtri(const double* Ls, int64_t lds, int64_t sL, double* X, int64_t ldx, int64_t sX,
                int m, int batch, double* W, hipStream_t s, bool halving = false,
                bool zero_upper = true) {
  if (m <= 0 || batch <= 0) return ROCBTA_OK;
  const int nt = cdiv(m, TT);
  const int nfull = m / TT;
  const int tail = m - nfull * TT;
  if (ldx != m || sX != (int64_t)m * m) return ROCBTA_ERR_ARG;  // X is a dense stack
  const bool halves = halving && tail == 0 && (nt & (nt - 1)) == 0;
  // 128-tiles clip the K range of a triangular operand at 128, which reaches
  // into the strictly upper 64-tiles: those must then be zero
  const bool skip = !zero_upper && halves && !gemm_tile128_enabled();
  int rc = skip ? ROCBTA_OK : fill(X, (int64_t)batch * m * m, 0.0, s);
  if (rc) return rc;
  rc = tile_factor(1, const_cast<double*>(Ls), lds, (int64_t)TT * (lds + 1), sL, X, ldx,
                   (int64_t)TT * (ldx + 1), sX, TT, nfull, batch, nullptr, 0, 0, 0, 0, s);
  if (rc) return rc;
  if (tail) {
    rc = tile_factor(1, const_cast<double*>(Ls) + (int64_t)nfull * TT * (lds + 1), lds, 0, sL,
                     X + (int64_t)nfull * TT * (ldx + 1), ldx, 0, sX, tail, 1, batch, nullptr, 0,
                     0, 0, 0, s);
    if (rc) return rc;
  }
  if (halves) {
    // Recursive halving, bottom-up: with L = [A 0; C B] (halves of order h),
    //   L^{-1} = [A^{-1} 0; -B^{-1} (C A^{-1}) B^{-1}],
    // so level h forms W = C A^{-1}, then X21 = -B^{-1} W, for all m / 2h
    // pairs of every matrix in two batched launches.  log2(nt) levels of
    // large products replace nt - 1 levels of 64-wide ones.  W holds m^2 / 4
    // doubles per matrix.
    const int64_t wstride = (int64_t)m * m / 4;
    for (int h = TT; h < m; h *= 2) {
      const int npair = m / (2 * h);
      const int64_t pL = (int64_t)2 * h * (lds + 1), pX = (int64_t)2 * h * (ldx + 1);
      GemmList w, x;
      push_term(w.add(W, h, h, h, 0.0, TRI_NONE, npair, (int64_t)h * h, batch, wstride),
                term(Ls + (int64_t)h * lds, lds, false, X, ldx, false, h, 1.0, TRI_NONE, TRI_LOWER,
                     pL, pX, sL, sX));
      push_term(x.add(X + (int64_t)h * ldx, ldx, h, h, 0.0, TRI_NONE, npair, pX, batch, sX),
                term(X + (int64_t)h * (ldx + 1), ldx, false, W, h, false, h, -1.0, TRI_LOWER,
                     TRI_NONE, pX, (int64_t)h * h, sX, wstride));
      rc = run_gemms(w, s);
      if (rc) return rc;
      rc = run_gemms(x, s);
      if (rc) return rc;
    }
    return ROCBTA_OK;
  }
  const int64_t wstride = (int64_t)nt * TT * TT;
  for (int d = 1; d < nt; ++d) {
    // j = 0 .. nt-1-d ; row tile r = j + d.  Split off the partial last row tile.
    const int nj = nt - d;
    const int nj_full = tail ? nj - 1 : nj;
    GemmList w, x;
    auto add_range = [&](int j0, int cnt, int Mr) {
      if (cnt <= 0) return;
      const int r0 = j0 + d;
      const double* A = Ls + (int64_t)r0 * TT * lds + (int64_t)j0 * TT;
      const double* Bx = X + (int64_t)j0 * TT * ldx + (int64_t)j0 * TT;
      double* Wc = W + (int64_t)j0 * TT * TT;
      push_term(w.add(Wc, TT, Mr, TT, 0.0, TRI_NONE, cnt, TT * TT, batch, wstride),
                term(A, lds, false, Bx, ldx, false, d * TT, 1.0, TRI_NONE, TRI_LOWER,
                     (int64_t)TT * (lds + 1), (int64_t)TT * (ldx + 1), sL, sX));
      double* Xc = X + (int64_t)r0 * TT * ldx + (int64_t)j0 * TT;
      const double* Xd = X + (int64_t)r0 * TT * ldx + (int64_t)r0 * TT;
      push_term(x.add(Xc, ldx, Mr, TT, 0.0, TRI_NONE, cnt, (int64_t)TT * (ldx + 1), batch, sX),
                term(Xd, ldx, false, Wc, TT, false, Mr, -1.0, TRI_LOWER, TRI_NONE,
                     (int64_t)TT * (ldx + 1), TT * TT, sX, wstride));
    };
    add_range(0, nj_full, TT);
    if (tail) add_range(nj - 1, 1, tail);
    rc = run_gemms(w, s);
    if (rc) return rc;
    rc = run_gemms(x, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

// -------------------------------------------------------- workspace layout
// Side streams (default priority) with growable event pools, used to
// overlap independent precompute with the caller's stream.  A call checks
// one out of its device's pool for its whole duration and returns it at the
// end (SideLease), so concurrent calls from several host threads never
// record or wait on each other's events.  Reuse after a call is safe: a
// stream wait binds to the event record that precedes it at enqueue time.
constexpr int SI_CHUNK = 16;
struct SideStream {
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> ev;
};

struct SidePool {
  std::mutex mu;
  std::vector<SideStream*> free_list[64];
};
SidePool& side_pool() {
  static SidePool pool;  // streams and events live until the process ends
  return pool;
}

class SideLease {
 public:
  SideLease(int nevents) {
    if (hipGetDevice(&dev_) != hipSuccess || dev_ < 0 || dev_ >= 64) return;
    SidePool& pool = side_pool();
    {
      std::lock_guard<std::mutex> lock(pool.mu);
      if (!pool.free_list[dev_].empty()) {
        ss_ = pool.free_list[dev_].back();
        pool.free_list[dev_].pop_back();
      }
    }
    if (!ss_) {
      ss_ = new SideStream();
      if (hipStreamCreateWithFlags(&ss_->s, hipStreamNonBlocking) != hipSuccess) {
        delete ss_;
        ss_ = nullptr;
        return;
      }
    }
    while ((int)ss_->ev.size() < nevents) {
      hipEvent_t e;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        release();
        return;
      }
      ss_->ev.push_back(e);
    }
  }
  ~SideLease() { release(); }
  SideStream* get() const { return ss_; }

 private:
  void release() {
    if (!ss_) return;
    SidePool& pool = side_pool();
    std::lock_guard<std::mutex> lock(pool.mu);
    pool.free_list[dev_].push_back(ss_);
    ss_ = nullptr;
  }
  int dev_ = -1;
  SideStream* ss_ = nullptr;
};


struct SiLayout {
  size_t linv, g, gt, w, tinv, wt, part, cnt, total;
};

SiLayout si_layout(int64_t n, int64_t b, int64_t a) {
  SiLayout l{};
  size_t off = 0;
  auto take = [&](size_t doubles) {
    size_t o = off;
    off = align256(off + doubles * sizeof(double));
    return o;
  };
  const int64_t ntb = (b + TT - 1) / TT, nta = (a + TT - 1) / TT;
  const int64_t nch = std::min<int64_t>(n, SI_CHUNK);  // L_ii^{-1} of one chunk at a time
  l.linv = take((size_t)nch * b * b);
  l.g = take((size_t)std::max<int64_t>(n - 1, 0) * b * b);
  l.gt = take((size_t)n * a * b);
  // trtri scratch: 64-wide levels need ntb * 64^2 per block, halving b^2 / 4
  l.w = take((size_t)nch * std::max<int64_t>(ntb * TT * TT, ntb * TT * ntb * TT / 4));
  l.tinv = take((size_t)a * a);
  l.wt = take((size_t)nta * TT * TT);
  // split-K partial tiles: up to 2 splits of 64-tiles or 8 splits of 128-tiles
  const int64_t t128 = (b + 127) / 128, a128 = (a + 127) / 128;
  l.part = take((size_t)std::max<int64_t>((ntb * ntb + ntb * nta + ntb) * 2 * TT * TT,
                                          (t128 * t128 + t128 * a128 + t128) * 8 * 128 * 128));
  l.cnt = take((size_t)(ntb * ntb + ntb + 64) / 2 + 1);  // split-K tile counters (ints)
  l.total = off;
  return l;
}

struct AfLayout {
  size_t tinv, counter, claim, chain_cu, hflags, cflags, linv, tipp, save_d, save_r, save_t, total;
  int nt, NT, slots, nstages, ntasks;
};

AfLayout af_layout(int64_t n, int64_t b, int64_t a) {
  AfLayout l{};
  l.nt = (int)((b + CT - 1) / CT);
  l.NT = (int)(n * l.nt);
  l.slots = 5 + 2 * l.nt;
  l.nstages = l.NT + 2;
  l.ntasks = l.nstages * l.slots;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  l.tinv = take(TT * TT * sizeof(double));
  l.counter = take(sizeof(int));
  l.claim = take(sizeof(int));
  l.chain_cu = take(sizeof(int));
  l.hflags = take((size_t)l.ntasks * sizeof(int));
  l.cflags = take((size_t)2 * l.NT * sizeof(int));
  l.linv = take((size_t)l.NT * CT * CT * sizeof(double));
  l.tipp = take((size_t)n * a * a * sizeof(double));
  // factorize_last_block = 0 on the chain: the last block, its arrow block
  // and the tip as given
  l.save_d = take((size_t)b * b * sizeof(double));
  l.save_r = take((size_t)a * b * sizeof(double));
  l.save_t = take((size_t)a * a * sizeof(double));
  l.total = off;
  return l;
}

int device_cus() {
  static std::atomic<int> cus[64];  // zero-initialised (static storage)
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int v = cus[dev].load(std::memory_order_relaxed);
  if (v == 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      v = 256;
    cus[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// ROCBTA_POBTAF_ALGO=loop forces the launch-per-tile path (used by tests to
// cover both paths); default: the persistent chain kernel whenever it applies.
bool pobtaf_use_chain(int64_t a) {
  const char* e = getenv("ROCBTA_POBTAF_ALGO");
  if (e && strcmp(e, "loop") == 0) return false;
  return a <= CT;
}

// info = 0 where it names a row of the last block or of the tip (rows past
// `first_open`), which factorize_last_block = 0 leaves unfactored.
__global__ void clear_info_above_kernel(int* info, int64_t first_open) {
  if (threadIdx.x == 0 && *info > first_open) *info = 0;
}

// T = T0 - sum_{i < nblk} tipp[i], summed in block order (the order the
// chain's tip task uses).  The chain forms only the lower part of each
// symmetric tipp[i] (its upper-right 32 x 32 quadrant stays zero), so both
// (r, s) and (s, r) take the lower entry.
__global__ void tip_open_kernel(double* T, const double* T0, const double* tipp, int a, int nblk) {
  const int aa = a * a;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < aa; idx += gridDim.x * blockDim.x) {
    const int r = idx / a, c = idx % a;
    const int lo = r >= c ? idx : c * a + r;
    double t = T0[idx];
    for (int i = 0; i < nblk; ++i) t -= tipp[(int64_t)i * aa + lo];
    T[idx] = t;
  }
}

PobtafChain chain_params(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                         double* T, int* info, char* ws, const AfLayout& lay) {
  PobtafChain P{};
  P.D = D;
  P.L = L;
  P.R = R;
  P.T = T;
  P.n = (int)n;
  P.b = (int)b;
  P.a = (int)a;
  P.nt = lay.nt;
  P.ha = a > 0 ? 1 : 0;
  P.NT = lay.NT;
  P.slots = lay.slots;
  P.nstages = lay.nstages;
  P.ntasks = lay.ntasks;
  if (ws) {
    P.counter = (int*)(ws + lay.counter);
    P.claim = (int*)(ws + lay.claim);
    P.chain_cu = (int*)(ws + lay.chain_cu);
    P.hflags = (int*)(ws + lay.hflags);
    P.cflags = (int*)(ws + lay.cflags);
    P.linv = (double*)(ws + lay.linv);
    P.tipp = (double*)(ws + lay.tipp);
  }
  P.info = info;
  P.spin_limit = 1ll << 22;
  P.f32 = g_mixed ? 1 : 0;
  // ROCBTA_CHAIN=classic keeps the step-after-step chain schedule
  const char* e = getenv("ROCBTA_CHAIN");
  const bool classic = e && strcmp(e, "classic") == 0;
  P.fused = (!classic && b % CT == 0) ? 1 : 0;
  P.trace = nullptr;
  return P;
}

__global__ void fill_u64_kernel(unsigned long long* p, int64_t count, unsigned long long v);

// Diagnostic: with ROCBTA_CHAIN_TRACE=<file> set, a single-matrix chain
// launch records its step timeline (chain_trace events) and appends it to
// <file> as text after the launch (synchronising the stream).
long long* chain_trace_buffer(size_t count) {
  static long long* buf = nullptr;
  static size_t cap = 0;
  if (count > cap) {
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    if (hipMalloc(&buf, count * sizeof(long long)) != hipSuccess) return nullptr;
    cap = count;
  }
  return buf;
}

int pobtaf_chain(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                 int* info, char* ws, const AfLayout& lay, hipStream_t s) {
  PobtafChain P = chain_params(n, b, a, D, L, R, T, info, ws, lay);
  const int64_t nflag = (int64_t)((lay.linv - lay.counter) / sizeof(int));
  fill_i32_kernel<<<(int)std::min<int64_t>(1024, (nflag + 255) / 256), 256, 0, s>>>(P.counter, nflag, 0);
  RB_CHECK(hipGetLastError());
  const char* trace_path = getenv("ROCBTA_CHAIN_TRACE");
  const size_t nev = (size_t)CHAIN_TRACE_EV * P.NT;
  if (trace_path) {
    P.trace = chain_trace_buffer(nev);
    if (P.trace) {
      fill_u64_kernel<<<64, 256, 0, s>>>((unsigned long long*)P.trace, (int64_t)nev, 0ull);
      RB_CHECK(hipGetLastError());
    }
  }
  const int grid = std::max(2, std::min(P.ntasks + 1, 2 * device_cus()));
  pobtaf_chain_kernel<<<grid, CH_THREADS, 0, s>>>(P);
  RB_CHECK(hipGetLastError());
  if (P.trace) {
    std::vector<long long> h(nev);
    RB_CHECK(hipStreamSynchronize(s));
    RB_CHECK(hipMemcpy(h.data(), P.trace, nev * sizeof(long long), hipMemcpyDeviceToHost));
    if (FILE* f = fopen(trace_path, "a")) {
      fprintf(f, "# chain trace n=%lld b=%lld a=%lld NT=%d fused=%d\n", (long long)n, (long long)b,
              (long long)a, P.NT, P.fused);
      for (int g = 0; g < P.NT; ++g) {
        for (int e = 0; e < CHAIN_TRACE_EV; ++e) fprintf(f, "%lld ", h[(size_t)CHAIN_TRACE_EV * g + e]);
        fprintf(f, "\n");
      }
      fclose(f);
    }
  }
  return ROCBTA_OK;
}

// ------------------------------------------------ persistent solve sweep
struct SwLayout {
  size_t status, claim, counter, tinv, ybuf, part, apart, total;
  int nt, ta, NT, NG, rc, nchunks, ntasks;
};

int sweep_rc(int64_t r) { return r <= 1 ? 1 : (r <= 4 ? 4 : 8); }

// At most this many column chunks (chain workgroups) per sweep launch.  A
// chain waits for helper partial sums, so chains may take only a fraction of
// the resident workgroup slots (two sweep workgroups fit per CU); wider
// right-hand sides run as several launches over groups of columns.
constexpr int SW_MAX_CHAINS = 64;

SwLayout sw_layout(int64_t n, int64_t b, int64_t a, int64_t r) {
  SwLayout l{};
  l.nt = cdiv(b, SW_T);
  l.ta = cdiv(a, SW_T);
  l.NT = (int)(n * l.nt);
  l.NG = l.NT + l.ta;
  l.rc = sweep_rc(r);
  l.nchunks = std::min(SW_MAX_CHAINS, std::max(1, cdiv(r, l.rc)));  // per launch
  l.ntasks = 2 * l.NT * l.nchunks;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  // status first: the caller may read it at workspace offset 0
  l.status = take(sizeof(int));
  l.claim = take(sizeof(int));
  l.counter = take(sizeof(int));
  l.tinv = take((size_t)l.NG * SW_T * SW_T * sizeof(double));
  // published blocks (pre-filled with the empty marker), contiguous
  l.ybuf = take((size_t)l.nchunks * l.NG * SW_T * l.rc * sizeof(double));
  l.part = take((size_t)l.nchunks * l.NG * SW_T * l.rc * sizeof(double));
  l.apart = take((size_t)l.nchunks * n * l.ta * SW_T * l.rc * sizeof(double));
  l.total = off;
  return l;
}

__global__ void fill_u64_kernel(unsigned long long* p, int64_t count, unsigned long long v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

// Solve algorithm: the block-level sweep (precomputed diagonal-block
// inverses, two dependent device-wide products per block row) for large
// blocks and at most 8 right-hand sides, its wide form (64 columns per
// launch, MFMA tiles) for more, the 64-row-tile sweep otherwise, the
// launch-per-tile form for partial solves on small blocks.
// ROCBTA_POBTAS_ALGO = block | wide | sweep | loop forces one (tests cover
// all four; block and wide fall back when the shape does not fit them).
enum SolveAlgo { SOLVE_LOOP = 0, SOLVE_SWEEP = 1, SOLVE_BLOCK = 2, SOLVE_WIDE = 3 };
constexpr int BS_MIN_B = 256;   // default: block sweep from this block size on
constexpr int BS_MAX_G = 1024;  // workgroups of the block sweep (tip partial sums sized for it)

int bs_cw(int64_t r) { return r <= 1 ? 1 : (r <= 4 ? 4 : 8); }
bool bs_fits(int64_t a, int64_t r) { return a * bs_cw(r) <= BS_TIPMAX; }
// Wider right-hand sides take the 64-row-tile sweep by default: the block
// sweep re-reads the factor once per group of 8 columns (measured at n=256,
// b=1024, a=32, N+C: r=8 19.8 ms, r=64 89 ms against 59 and 58 ms for the
// sweep).
constexpr int BS_MAX_R = 8;

// The wide sweep: full solves with b a multiple of its 16-row tiles.
bool wd_fits(int64_t b, int partial) { return !partial && b >= WD_CW && b % WD_T == 0; }

int pobtas_algo(int64_t b, int64_t a, int64_t r, int partial) {
  const char* e = getenv("ROCBTA_POBTAS_ALGO");
  if (e && strcmp(e, "loop") == 0) return SOLVE_LOOP;
  if (e && strcmp(e, "wide") == 0)
    return wd_fits(b, partial) ? SOLVE_WIDE : (partial ? SOLVE_LOOP : SOLVE_SWEEP);
  if (e && strcmp(e, "sweep") == 0) return partial ? SOLVE_LOOP : SOLVE_SWEEP;
  if (e && strcmp(e, "block") == 0) return bs_fits(a, r) ? SOLVE_BLOCK : (partial ? SOLVE_LOOP : SOLVE_SWEEP);
  if (b >= BS_MIN_B && r <= BS_MAX_R && bs_fits(a, r)) return SOLVE_BLOCK;
  if (b >= BS_MIN_B && r > BS_MAX_R && wd_fits(b, partial)) return SOLVE_WIDE;
  return partial ? SOLVE_LOOP : SOLVE_SWEEP;
}

// Inverses of the diagonal tiles of an m_total x m_total lower-triangular
// matrix stored with row stride ld, for `batch` matrices (stride sA), into
// consecutive 64 x 64 tiles starting at X (per matrix: ceil(m_total/64) tiles).
int diag_tile_inverses(const double* A, int64_t ld, int64_t sA, int m_total, int batch, double* X,
                       int64_t sX, hipStream_t s) {
  const int nfull = m_total / SW_T, rem = m_total % SW_T;
  double* Am = const_cast<double*>(A);
  if (nfull > 0) {
    int rc = tile_factor(1, Am, ld, sA, (int64_t)SW_T * (ld + 1), X, SW_T, sX, SW_T * SW_T, SW_T,
                         batch, nfull, nullptr, 0, 0, 0, 0, s);
    if (rc) return rc;
  }
  if (rem > 0) {
    int rc = tile_factor(1, Am + (int64_t)nfull * SW_T * (ld + 1), ld, sA, 0,
                         X + (int64_t)nfull * SW_T * SW_T, SW_T, sX, 0, rem, batch, 1, nullptr,
                         0, 0, 0, 0, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

template <int RC>
int launch_sweep(const SweepP& P, int grid, hipStream_t s) {
  pobtas_sweep_kernel<RC><<<grid, SW_THREADS, 0, s>>>(P);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

int pobtas_sweep(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                 const double* R, const double* T, double* B, int64_t r, int64_t ldb, int trans,
                 char* ws, hipStream_t s) {
  const SwLayout lay = sw_layout(n, b, a, r);
  SweepP P{};
  P.D = D;
  P.L = L;
  P.R = R;
  P.T = T;
  P.B = B;
  P.ldb = ldb;
  P.n = (int)n;
  P.b = (int)b;
  P.a = (int)a;
  P.nt = lay.nt;
  P.ta = lay.ta;
  P.NT = lay.NT;
  P.NG = lay.NG;
  P.r = (int)r;
  P.rc = lay.rc;
  P.nchunks = lay.nchunks;
  P.trans = trans;
  P.tinv = (const double*)(ws + lay.tinv);
  P.ybuf = (double*)(ws + lay.ybuf);
  P.part = (double*)(ws + lay.part);
  P.apart = (double*)(ws + lay.apart);
  P.claim = (int*)(ws + lay.claim);
  P.counter = (int*)(ws + lay.counter);
  P.status = (int*)(ws + lay.status);
  P.ntasks = lay.ntasks;
  P.spin_limit = 1ll << 22;
  double* tinv = (double*)(ws + lay.tinv);
  int rc = diag_tile_inverses(D, b, b * b, (int)b, (int)n, tinv, (int64_t)lay.nt * SW_T * SW_T, s);
  if (rc) return rc;
  if (a > 0) {
    rc = diag_tile_inverses(T, a, 0, (int)a, 1, tinv + (int64_t)lay.NT * SW_T * SW_T, 0, s);
    if (rc) return rc;
  }
  set_int_kernel<<<1, 64, 0, s>>>(P.status, 0);  // kept across the column groups
  RB_CHECK(hipGetLastError());
  const int64_t group_cols = (int64_t)lay.nchunks * lay.rc;
  const int64_t nempty = (int64_t)((lay.total - lay.ybuf) / sizeof(double));
  for (int64_t c0 = 0; c0 < r; c0 += group_cols) {
    P.B = B + c0;
    P.r = (int)std::min<int64_t>(group_cols, r - c0);
    P.nchunks = cdiv(P.r, lay.rc);
    P.ntasks = 2 * lay.NT * P.nchunks;
    // claim ticket and helper queue head (the slots between them included)
    fill_i32_kernel<<<1, 256, 0, s>>>(P.claim, (int64_t)((lay.tinv - lay.claim) / sizeof(int)), 0);
    RB_CHECK(hipGetLastError());
    fill_u64_kernel<<<(int)std::min<int64_t>(4096, (nempty + 255) / 256), 256, 0, s>>>(
        (unsigned long long*)(ws + lay.ybuf), nempty, SW_EMPTY);
    RB_CHECK(hipGetLastError());
    const int grid = std::max(P.nchunks + 1, std::min(P.nchunks + P.ntasks, 2 * device_cus()));
    switch (lay.rc) {
      case 1:
        rc = launch_sweep<1>(P, grid, s);
        break;
      case 4:
        rc = launch_sweep<4>(P, grid, s);
        break;
      default:
        rc = launch_sweep<8>(P, grid, s);
        break;
    }
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

struct BsLayout {
  size_t status, linv, lt, w, tinv, wt, sbuf, tbuf, tpart, tipbuf, total;
  int cw;
};

BsLayout bs_layout(int64_t n, int64_t b, int64_t a, int64_t r) {
  BsLayout l{};
  l.cw = bs_cw(r);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const int64_t ntb = (b + TT - 1) / TT, nta = (a + TT - 1) / TT;
  l.status = take(sizeof(int));
  l.linv = take((size_t)n * b * b * sizeof(double));
  l.lt = take((size_t)std::max<int64_t>(n - 1, 1) * b * b * sizeof(double));  // backward: L^T
  // trtri scratch: 64-wide levels need ntb * 64^2 per block, halving b^2 / 4
  l.w = take((size_t)n * std::max<int64_t>(ntb * TT * TT, ntb * TT * ntb * TT / 4) * sizeof(double));
  l.tinv = take((size_t)a * a * sizeof(double));
  l.wt = take((size_t)nta * TT * TT * sizeof(double));
  l.sbuf = take((size_t)n * b * l.cw * sizeof(double));
  l.tbuf = take((size_t)n * b * l.cw * sizeof(double));
  l.tpart = take((size_t)BS_MAX_G * a * l.cw * sizeof(double));
  l.tipbuf = take((size_t)a * l.cw * sizeof(double));
  l.total = off;
  return l;
}

template <int CW>
int launch_block_sweep(BlockSweepP P, hipStream_t s) {
  const int rw = bs_rw(CW, P.trans);
  P.G = (int)std::min<int64_t>(std::min<int64_t>((P.b + rw - 1) / rw, BS_MAX_G), device_cus());
  if (P.trans == 0)
    pobtas_block_kernel<CW, 0><<<P.G, BS_THREADS, 0, s>>>(P);
  else
    pobtas_block_kernel<CW, 1><<<P.G, BS_THREADS, 0, s>>>(P);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

// pobtas with the block-level sweep: the inverses of the diagonal blocks and
// of the tip factor first (batched trtri), then one persistent launch per
// group of at most CW right-hand-side columns.
int pobtas_block(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                 const double* R, const double* T, double* B, int64_t r, int64_t ldb, int trans,
                 int partial, char* ws, hipStream_t s) {
  const BsLayout lay = bs_layout(n, b, a, r);
  int* status = (int*)(ws + lay.status);
  set_int_kernel<<<1, 64, 0, s>>>(status, 0);
  RB_CHECK(hipGetLastError());
  double* Linv = (double*)(ws + lay.linv);
  double* Tinv = (double*)(ws + lay.tinv);
  // Linv is read on and below the diagonal tiles only (the backward sweep
  // reads it after the in-place transpose, on and above them)
  int rc = block_trtri(D, b, b * b, Linv, b, b * b, (int)b, (int)n, (double*)(ws + lay.w), s, true,
                       false);
  if (rc) return rc;
  if (a > 0) {
    rc = block_trtri(T, a, 0, Tinv, a, a * a, (int)a, 1, (double*)(ws + lay.wt), s);
    if (rc) return rc;
  }
  if (trans) {
    // the backward sweep reads L_{i+1,i}^T and (the upper triangle of)
    // Linv_i^T by rows
    const int T = (int)((b + 63) / 64);
    for (int64_t z0 = 0; z0 < n; z0 += 65535) {
      const int nz = (int)std::min<int64_t>(65535, n - z0);
      transpose_lower_to_upper_kernel<<<dim3(T * (T + 1) / 2, nz), 256, 0, s>>>(
          Linv + z0 * b * b, (int)b, b * b);
      RB_CHECK(hipGetLastError());
    }
    for (int64_t z0 = 0; z0 + 1 < n; z0 += 65535) {
      const int nz = (int)std::min<int64_t>(65535, n - 1 - z0);
      transpose_blocks_kernel<<<dim3(T * T, nz), 256, 0, s>>>(L + z0 * b * b,
                                                             (double*)(ws + lay.lt) + z0 * b * b,
                                                             (int)b, b * b, T);
      RB_CHECK(hipGetLastError());
    }
  }
  BlockSweepP P{};
  P.L = trans ? (const double*)(ws + lay.lt) : L;
  P.R = R;
  P.Linv = Linv;
  P.Tinv = Tinv;
  P.ldb = ldb;
  P.n = (int)n;
  P.b = (int)b;
  P.a = (int)a;
  P.trans = trans;
  P.partial = partial;
  P.sbuf = (double*)(ws + lay.sbuf);
  P.tbuf = (double*)(ws + lay.tbuf);
  P.tpart = (double*)(ws + lay.tpart);
  P.tipbuf = (double*)(ws + lay.tipbuf);
  P.status = status;
  P.spin_limit = 1ll << 22;
  const int64_t nempty = (int64_t)((lay.total - lay.sbuf) / sizeof(double));
  for (int64_t c0 = 0; c0 < r; c0 += lay.cw) {
    P.B = B + c0;
    P.r = (int)std::min<int64_t>(lay.cw, r - c0);
    fill_u64_kernel<<<(int)std::min<int64_t>(4096, (nempty + 255) / 256), 256, 0, s>>>(
        (unsigned long long*)(ws + lay.sbuf), nempty, SW_EMPTY);
    RB_CHECK(hipGetLastError());
    switch (lay.cw) {
      case 1:
        rc = launch_block_sweep<1>(P, s);
        break;
      case 4:
        rc = launch_block_sweep<4>(P, s);
        break;
      default:
        rc = launch_block_sweep<8>(P, s);
        break;
    }
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

// Workspace of the wide sweep: the status word first (read by the caller),
// the block inverses as for the block sweep, the published rows of one
// 64-column launch, and the tip's per-block products.
struct WdLayout {
  size_t status, linv, w, tinv, wt, sbuf, tbuf, tipp, tipb, total;
};

WdLayout wd_layout(int64_t n, int64_t b, int64_t a) {
  WdLayout l{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const int64_t ntb = (b + TT - 1) / TT, nta = (a + TT - 1) / TT;
  l.status = take(sizeof(int));
  l.linv = take((size_t)n * b * b * sizeof(double));
  l.w = take((size_t)n * std::max<int64_t>(ntb * TT * TT, ntb * TT * ntb * TT / 4) * sizeof(double));
  l.tinv = take((size_t)a * a * sizeof(double));
  l.wt = take((size_t)nta * TT * TT * sizeof(double));
  l.sbuf = take((size_t)n * b * WD_CW * sizeof(double));
  l.tbuf = take((size_t)n * b * WD_CW * sizeof(double));
  l.tipp = take((size_t)n * a * WD_CW * sizeof(double));
  l.tipb = take((size_t)a * WD_CW * sizeof(double));
  l.total = off;
  return l;
}

// pobtas with the wide sweep, per group of at most 64 right-hand-side
// columns.  Forward: the sweep gives y_0 .. y_{n-1}; then the tip, y_t =
// Tinv (b_t - sum_i R_i y_i) (per-block products, a fixed-order sum, one
// product).  Backward: x_t = Tinv^T y_t and y'_i = y_i - R_i^T x_t (one
// batched product) first, then the sweep.
int pobtas_wide(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                const double* R, const double* T, double* B, int64_t r, int64_t ldb, int trans,
                char* ws, hipStream_t s) {
  const WdLayout lay = wd_layout(n, b, a);
  int* status = (int*)(ws + lay.status);
  set_int_kernel<<<1, 64, 0, s>>>(status, 0);
  RB_CHECK(hipGetLastError());
  double* Linv = (double*)(ws + lay.linv);
  double* Tinv = (double*)(ws + lay.tinv);
  double* tipp = (double*)(ws + lay.tipp);
  double* tipb = (double*)(ws + lay.tipb);
  // Linv is read on and below the diagonal tiles only
  int rc = block_trtri(D, b, b * b, Linv, b, b * b, (int)b, (int)n, (double*)(ws + lay.w), s, true,
                       false);
  if (rc) return rc;
  if (a > 0) {
    rc = block_trtri(T, a, 0, Tinv, a, a * a, (int)a, 1, (double*)(ws + lay.wt), s);
    if (rc) return rc;
  }
  WideSweepP P{};
  P.L = L;
  P.Linv = Linv;
  P.ldb = ldb;
  P.n = (int)n;
  P.b = (int)b;
  P.trans = trans;
  P.sbuf = (double*)(ws + lay.sbuf);
  P.tbuf = (double*)(ws + lay.tbuf);
  P.status = status;
  P.spin_limit = 1ll << 22;
  const int64_t nempty = (int64_t)n * b * WD_CW * 2;  // sbuf and tbuf (adjacent)
  const int ai = (int)a, bi = (int)b;
  const int64_t ab = a * b;
  double* Bt = B + n * b * ldb;  // tip rows
  for (int64_t c0 = 0; c0 < r; c0 += WD_CW) {
    const int rc_ = (int)std::min<int64_t>(WD_CW, r - c0);
    if (trans && a > 0) {
      // x_t = Tinv^T y_t, then y'_i = y_i - R_i^T x_t for every block
      wide_copy_rows_kernel<<<cdiv(ai * rc_, 256), 256, 0, s>>>(tipb, rc_, Bt + c0, ldb, ai, rc_);
      RB_CHECK(hipGetLastError());
      GemmList xt;
      push_term(xt.add(Bt + c0, ldb, ai, rc_, 0.0),
                term(Tinv, a, true, tipb, rc_, false, ai, 1.0, TRI_UPPER, TRI_NONE));
      if ((rc = run_gemms(xt, s))) return rc;
      GemmList ry;
      push_term(ry.add(B + c0, ldb, bi, rc_, 1.0, TRI_NONE, (int)n, b * ldb),
                term(R, b, true, Bt + c0, ldb, false, ai, -1.0, TRI_NONE, TRI_NONE, ab, 0));
      if ((rc = run_gemms(ry, s))) return rc;
    }
    P.B = B + c0;
    P.r = rc_;
    fill_u64_kernel<<<(int)std::min<int64_t>(4096, (nempty + 255) / 256), 256, 0, s>>>(
        (unsigned long long*)P.sbuf, nempty, SW_EMPTY);
    RB_CHECK(hipGetLastError());
    const int nitems = (bi / WD_T) * ((rc_ + WD_T - 1) / WD_T);
    P.G = std::min(std::min(nitems, WD_MAX_G), device_cus());
    if (trans)
      pobtas_wide_kernel<1><<<P.G, WD_THREADS, 0, s>>>(P);
    else
      pobtas_wide_kernel<0><<<P.G, WD_THREADS, 0, s>>>(P);
    RB_CHECK(hipGetLastError());
    if (!trans && a > 0) {
      // y_t = Tinv (b_t - sum_i R_i y_i)
      GemmList pr;
      push_term(pr.add(tipp, rc_, ai, rc_, 0.0, TRI_NONE, (int)n, ai * rc_),
                term(R, b, false, B + c0, ldb, false, bi, 1.0, TRI_NONE, TRI_NONE, ab, b * ldb));
      if ((rc = run_gemms(pr, s))) return rc;
      wide_tip_reduce_kernel<<<cdiv(ai * rc_, 256), 256, 0, s>>>(tipb, Bt + c0, ldb, tipp, ai, rc_,
                                                               (int)n);
      RB_CHECK(hipGetLastError());
      GemmList yt;
      push_term(yt.add(Bt + c0, ldb, ai, rc_, 0.0),
                term(Tinv, a, false, tipb, rc_, false, ai, 1.0, TRI_LOWER, TRI_NONE));
      if ((rc = run_gemms(yt, s))) return rc;
    }
  }
  return ROCBTA_OK;
}

// Workspace of the permuted selected inversion: the inverses of the diagonal
// blocks (all n, by the batched trtri), its scratch, copies of the factor
// blocks a step overwrites while still reading them (L_{i+1,i}, L_{top,i},
// L_{t,i}), and the bracketed sums of the step's four products.
struct PsiLayout {
  size_t linv, w, lt, bt, rt, wl, wb, wr, wd, total;
};
PsiLayout psi_layout(int64_t n, int64_t b, int64_t a) {
  PsiLayout l{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const int64_t ntb = (b + TT - 1) / TT;
  const size_t bb = (size_t)b * b * sizeof(double), ab = (size_t)std::max<int64_t>(a, 1) * b * sizeof(double);
  l.linv = take((size_t)n * bb);
  l.w = take((size_t)n * std::max<int64_t>(ntb * TT * TT, ntb * TT * ntb * TT / 4) * sizeof(double));
  l.lt = take(bb);
  l.bt = take(bb);
  l.rt = take(ab);
  l.wl = take(bb);
  l.wb = take(bb);
  l.wr = take(ab);
  l.wd = take(bb);
  l.total = off;
  return l;
}

// Device workspace of the host-streaming forms: two slots of every block
// kind in flight (one being computed, one being copied), the tip, and the
// per-block scratch of the factorization / selected inversion.
struct StLayout {
  size_t d[2], l[2], r[2], xd[2], xl[2], xr[2], t, tinv, linv, w, g, gt, ti, wt, total;
};
StLayout st_layout(int64_t b, int64_t a, bool selinv) {
  StLayout l{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const size_t bb = (size_t)b * b * sizeof(double), ab = (size_t)std::max<int64_t>(a, 1) * b * sizeof(double);
  const size_t aa = (size_t)std::max<int64_t>(a, 1) * std::max<int64_t>(a, 1) * sizeof(double);
  for (int k = 0; k < 2; ++k) {
    l.d[k] = take(bb);
    l.l[k] = take(bb);
    l.r[k] = take(ab);
  }
  l.t = take(aa);
  l.tinv = take(TT * TT * sizeof(double));
  if (selinv) {
    for (int k = 0; k < 2; ++k) {
      l.xd[k] = take(bb);
      l.xl[k] = take(bb);
      l.xr[k] = take(ab);
    }
    const int64_t ntb = (b + TT - 1) / TT, nta = (std::max<int64_t>(a, 1) + TT - 1) / TT;
    l.linv = take(bb);
    l.w = take((size_t)std::max(ntb * TT * TT, ntb * TT * ntb * TT / 4) * sizeof(double));
    l.g = take(bb);
    l.gt = take(ab);
    l.ti = take(aa);
    l.wt = take((size_t)std::max(nta * TT * TT, nta * TT * nta * TT / 4) * sizeof(double));
  }
  l.total = off;
  return l;
}

// Workspace of the non-SPD sweeps: the blocked inverse's pivot-tile
// inverse P and its row / column panels, and the products of one step.
struct DdLayout {
  size_t p, r, c, w, v, b1, b2, c1, c2, e, g, dc, y;
  // quadratic (A X A^T = B) forms
  size_t bdc, df1, df2, dk1, dk2, sm, mt, q1, q2, q3, z, p1, p2, m1, m2, v1, zt, p3, d2, bbt, y2,
      total;
};
DdLayout dd_layout(int64_t b, int64_t a) {
  DdLayout l{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  const int64_t mm = std::max(b, a), a1 = std::max<int64_t>(a, 1);
  const size_t bb = (size_t)b * b * sizeof(double), ab = (size_t)a1 * b * sizeof(double);
  l.p = take(TT * TT * sizeof(double));
  l.r = take((size_t)TT * mm * sizeof(double));
  l.c = take((size_t)mm * TT * sizeof(double));
  l.w = take(bb);
  l.v = take(ab);
  l.b1 = take(bb);
  l.b2 = take(ab);
  l.c1 = take(bb);
  l.c2 = take(ab);
  l.e = take(bb);
  l.g = take(bb);
  l.dc = take(bb);
  l.y = take(ab);
  const size_t aa = (size_t)a1 * a1 * sizeof(double);
  l.bdc = take(bb);
  l.df1 = take(bb);
  l.df2 = take(ab);
  l.dk1 = take(bb);
  l.dk2 = take(ab);
  l.sm = take(bb);
  l.mt = take(bb);
  l.q1 = take(bb);
  l.q2 = take(bb);
  l.q3 = take(bb);
  l.z = take(bb);
  l.p1 = take(bb);
  l.p2 = take(ab);
  l.m1 = take(bb);
  l.m2 = take(ab);
  l.v1 = take(ab);
  l.zt = take(ab);
  l.p3 = take(ab);
  l.d2 = take(ab);
  l.bbt = take(std::max(ab, aa));
  l.y2 = take(ab);
  l.total = off;
  return l;
}

// A <- A^{-1} for a general m x m matrix (row stride ld), in place, by
// blocked Gauss-Jordan elimination over 64-wide pivot blocks without
// pivoting (the ddbt* blocks are diagonally dominant).  Step k, with
// P = A_kk^{-1}:
//   (1) A_kk <- P (tile kernel, in place);
//   (2) A_kj <- P A_kj (j != k);
//   (3) A_ij -= A_ik A_kj (i, j != k; A_kj is the new row block);
//   (4) A_ik <- -A_ik P (i != k).
// (2) and (4) are in place: the workgroup of an output tile reads, of the
// aliased operand, only that tile (K = the pivot block's width), and reads
// it in full before its epilogue writes.
int block_inverse(double* A, int64_t ld, int m, char* ws, const DdLayout& lay, hipStream_t s) {
  (void)ws;
  (void)lay;
  for (int k0 = 0; k0 < m; k0 += TT) {
    const int tk = std::min(TT, m - k0), k1 = k0 + tk;
    double* Pk = A + (int64_t)k0 * ld + k0;
    tile_getri_kernel<<<1, 256, 0, s>>>(Pk, ld, tk, Pk, ld);
    RB_CHECK(hipGetLastError());
    if (tk == m) continue;
    double* Rk = A + (int64_t)k0 * ld;  // row block k
    const int cols[2][2] = {{0, k0}, {k1, m}};
    GemmList g2;
    for (const auto& cc : cols) {
      const int c0 = cc[0], nc = cc[1] - cc[0];
      if (nc > 0)
        push_term(g2.add(Rk + c0, ld, tk, nc, 0.0), term(Pk, ld, false, Rk + c0, ld, false, tk, 1.0));
    }
    int rc = run_gemms(g2, s);
    if (rc) return rc;
    GemmList g3, g4;
    for (const auto& rr : cols) {
      const int r0 = rr[0], nr = rr[1] - rr[0];
      if (nr <= 0) continue;
      double* Cik = A + (int64_t)r0 * ld + k0;  // A_ik, rows r0 ..
      for (const auto& cc : cols) {
        const int c0 = cc[0], nc = cc[1] - cc[0];
        if (nc > 0)
          push_term(g3.add(A + (int64_t)r0 * ld + c0, ld, nr, nc, 1.0),
                    term(Cik, ld, false, Rk + c0, ld, false, tk, -1.0));
      }
      push_term(g4.add(Cik, ld, nr, tk, 0.0), term(Cik, ld, false, Pk, ld, false, tk, -1.0));
    }
    if ((rc = run_gemms(g3, s))) return rc;
    if ((rc = run_gemms(g4, s))) return rc;
  }
  return ROCBTA_OK;
}

// dst = src^T for one b x b block (64 x 64 tiles through LDS).
int transpose_block(const double* src, double* dst, int64_t b, hipStream_t s) {
  const int T = (int)((b + 63) / 64);
  transpose_blocks_kernel<<<dim3(T * T, 1), 256, 0, s>>>(src, dst, (int)b, b * b, T);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

int logdet_parts(int64_t n, int64_t b, int64_t a) {
  return (int)((n * b + a + LD_ROWS_PER_WG - 1) / LD_ROWS_PER_WG);
}

bool bad_shape(int64_t n, int64_t b, int64_t a) {
  return n < 1 || b < 1 || a < 0 || b > (1 << 20) || a > (1 << 20) || n > (1 << 24);
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

const char* rocbta_version(void) { return "rocbta 0.1.0 (gfx950, fp64 mfma)"; }

size_t rocbta_workspace_bytes(int op, int64_t n, int64_t b, int64_t a, int64_t r) {
  (void)r;
  if (bad_shape(n, b, a)) return 0;
  switch (op) {
    case ROCBTA_OP_POBTAF:
      return af_layout(n, b, a).total;
    case ROCBTA_OP_POBTASI:
      return si_layout(n, b, a).total;
    case ROCBTA_OP_POBTAS: {
      // enough for the algorithm the shape selects (partial or not)
      const int64_t rr = std::max<int64_t>(r, 1);
      size_t need = sw_layout(n, b, a, rr).total;
      if (pobtas_algo(b, a, rr, 0) == SOLVE_BLOCK || pobtas_algo(b, a, rr, 1) == SOLVE_BLOCK)
        need = std::max(need, bs_layout(n, b, a, rr).total);
      if (pobtas_algo(b, a, rr, 0) == SOLVE_WIDE) need = std::max(need, wd_layout(n, b, a).total);
      return need;
    }
    case ROCBTA_OP_LOGDET:
      return align256((size_t)logdet_parts(n, b, a) * sizeof(double));
    case ROCBTA_OP_POBTAF_PERMUTED:
      return align256((size_t)TT * TT * sizeof(double));
    case ROCBTA_OP_POBTASI_PERMUTED:
      return psi_layout(n, b, a).total;
    case ROCBTA_OP_POBTAS_PERMUTED:
      return 256;
    case ROCBTA_OP_DDBTASC:
    case ROCBTA_OP_DDBTASCI:
      return dd_layout(b, a).total;
    case ROCBTA_OP_POBTAF_STREAMING:
      return st_layout(b, a, false).total;
    case ROCBTA_OP_POBTASI_STREAMING:
      return st_layout(b, a, true).total;
    case ROCBTA_OP_RESIDUAL: {
      const int64_t rr = std::max<int64_t>(r, 1), rc = sweep_rc(rr);
      return align256((size_t)cdiv(rr, rc) * n * cdiv(a, SW_T) * SW_T * rc * sizeof(double)) + 256;
    }
    default:
      return 0;
  }
}

int rocbta_dpobtaf(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                   int factorize_last_block, int* info, void* workspace, size_t workspace_bytes,
                   void* stream) {
  if (bad_shape(n, b, a) || !D || (n > 1 && !L) || (a > 0 && (!R || !T)))
    return ROCBTA_ERR_ARG;
  if (!workspace || workspace_bytes < rocbta_workspace_bytes(ROCBTA_OP_POBTAF, n, b, a, 0))
    return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const AfLayout lay = af_layout(n, b, a);
  double* tinv = (double*)((char*)workspace + lay.tinv);
  if (!info) return ROCBTA_ERR_ARG;
  set_int_kernel<<<1, 64, 0, s>>>(info, 0);
  RB_CHECK(hipGetLastError());
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  if (!factorize_last_block && n == 1) return ROCBTA_OK;  // no block row to eliminate
  if (pobtaf_use_chain(a)) {
    char* ws = (char*)workspace;
    if (factorize_last_block) return pobtaf_chain(n, b, a, D, L, R, T, info, ws, lay, s);
    // pobtaf.py:143-159 leaves D_{n-1}, R_{n-1} and T Schur-updated but not
    // factored.  The chain factors them too; they are rebuilt afterwards
    // from the values given: D_{n-1} - L_{n-1,n-2} L_{n-1,n-2}^T, R_{n-1} -
    // L_{t,n-2} L_{n-1,n-2}^T and T - sum_{i<n-1} L_{t,i} L_{t,i}^T (the
    // chain's per-block tip products).  Pivots of those blocks do not count.
    double* Dl = D + (n - 1) * bb;
    double* Rl = R + (n - 1) * ab;
    double* sd = (double*)(ws + lay.save_d);
    double* sr = (double*)(ws + lay.save_r);
    double* st = (double*)(ws + lay.save_t);
    RB_CHECK(hipMemcpyAsync(sd, Dl, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (a > 0) {
      RB_CHECK(hipMemcpyAsync(sr, Rl, ab * sizeof(double), hipMemcpyDeviceToDevice, s));
      RB_CHECK(hipMemcpyAsync(st, T, a * a * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    int rc = pobtaf_chain(n, b, a, D, L, R, T, info, ws, lay, s);
    if (rc) return rc;
    clear_info_above_kernel<<<1, 64, 0, s>>>(info, (n - 1) * b);
    RB_CHECK(hipGetLastError());
    RB_CHECK(hipMemcpyAsync(Dl, sd, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (a > 0) RB_CHECK(hipMemcpyAsync(Rl, sr, ab * sizeof(double), hipMemcpyDeviceToDevice, s));
    const double* Lp = L + (n - 2) * bb;
    GemmList upd;
    push_term(upd.add(Dl, b, bi, bi, 1.0), term(Lp, b, false, Lp, b, true, bi, -1.0));
    if (a > 0)
      push_term(upd.add(Rl, b, ai, bi, 1.0), term(R + (n - 2) * ab, b, false, Lp, b, true, bi, -1.0));
    rc = run_gemms(upd, s);
    if (rc) return rc;
    if (a > 0) {
      tip_open_kernel<<<cdiv(a * a, 256), 256, 0, s>>>(T, st, (double*)(ws + lay.tipp), ai,
                                                      (int)(n - 1));
      RB_CHECK(hipGetLastError());
    }
    return ROCBTA_OK;
  }
  for (int64_t i = 0; i + 1 < n; ++i) {
    double* Di = D + i * bb;
    std::vector<Panel> panels = {{L + i * bb, bi, b}};
    if (a > 0) panels.push_back({R + i * ab, ai, b});
    int rc = factor_block_column(Di, b, bi, panels, i * b, info, tinv, s);
    if (rc) return rc;
    // Schur complement update of block row i+1 (pobtaf.py:143-159)
    const bool full_last = (i + 2 == n) && !factorize_last_block;
    GemmList upd;
    const double* Li = L + i * bb;
    push_term(upd.add(D + (i + 1) * bb, b, bi, bi, 1.0, full_last ? TRI_NONE : TRI_LOWER),
              term(Li, b, false, Li, b, true, bi, -1.0));
    if (a > 0) {
      push_term(upd.add(R + (i + 1) * ab, b, ai, bi, 1.0),
                term(R + i * ab, b, false, Li, b, true, bi, -1.0));
      push_term(upd.add(T, a, ai, ai, 1.0), term(R + i * ab, b, false, R + i * ab, b, true, bi, -1.0));
    }
    rc = run_gemms(upd, s);
    if (rc) return rc;
  }
  if (factorize_last_block) {  // pobtaf.py:161-183
    double* Dl = D + (n - 1) * bb;
    std::vector<Panel> panels;
    if (a > 0) panels.push_back({R + (n - 1) * ab, ai, b});
    int rc = factor_block_column(Dl, b, bi, panels, (n - 1) * b, info, tinv, s);
    if (rc) return rc;
    if (a > 0) {
      GemmList upd;
      const double* Rl = R + (n - 1) * ab;
      push_term(upd.add(T, a, ai, ai, 1.0), term(Rl, b, false, Rl, b, true, bi, -1.0));
      rc = run_gemms(upd, s);
      if (rc) return rc;
      rc = factor_block_column(T, a, ai, {}, n * b, info, tinv, s);
      if (rc) return rc;
    }
  }
  return ROCBTA_OK;
}

int rocbta_dpobtasi(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                    int invert_last_block, void* workspace, size_t workspace_bytes,
                    void* stream) {
  if (bad_shape(n, b, a) || !D || (n > 1 && !L) || (a > 0 && (!R || !T)))
    return ROCBTA_ERR_ARG;
  const SiLayout lay = si_layout(n, b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double* Linv = (double*)(ws + lay.linv);
  double* G = (double*)(ws + lay.g);
  double* Gt = (double*)(ws + lay.gt);
  double* W = (double*)(ws + lay.w);
  double* Tinv = (double*)(ws + lay.tinv);
  double* Wt = (double*)(ws + lay.wt);
  double* part = (double*)(ws + lay.part);
  int* cnt = (int*)(ws + lay.cnt);
  const int ntb = cdiv(b, TT);
  // Split the K loop of the recurrence's b x b products when one product
  // alone has too few output tiles to occupy every CU (128-tiles: one
  // workgroup per CU; 64-tiles: two).
  const int ts = gemm_ts_for((int)b, (int)b, (int)b);
  const int nts = cdiv(b, ts);
  const int cus = device_cus();
  // 64-tiles: no split (one 1024^2 product is 256 tiles, one per CU, and a
  // split through global partials measured slower than none).
  int ksplit = 1;
  if (ts == 128) ksplit = std::max(1, std::min(8, cus / (nts * nts)));
  if (ksplit > 1) {
    fill_i32_kernel<<<1, 256, 0, s>>>(cnt, ntb * ntb + ntb, 0);
    RB_CHECK(hipGetLastError());
  }
  // the arrow problem of a step has its own split-K bookkeeping
  double* part_r = part + (int64_t)nts * nts * ksplit * ts * ts;
  int* cnt_r = cnt + nts * nts;
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  // Blocks whose factor is still stored: all of them, or all but the last
  // when the caller supplies X_{n-1,n-1}, X_{t,n-1}, X_T (pobtasi.py:114).
  const int npre = (int)(invert_last_block ? n : n - 1);

  // (1)+(2) Independent per-block precompute, in chunks of blocks from the
  // last one down, on a side stream so that it overlaps the sequential
  // recurrence on the caller's stream:
  //   L_ii^{-1}; G_i = L_{i+1,i} L_ii^{-1}; Gt_i = L_{t,i} L_ii^{-1};
  //   S_i = L_ii^{-T} L_ii^{-1} -> D_i.
  const int nchunks = (npre + SI_CHUNK - 1) / SI_CHUNK;
  SideLease lease(nchunks + 1);
  SideStream* side = lease.get();
  if (!side) return ROCBTA_ERR_HIP;
  // ROCBTA_SI_OVERLAP=0 runs the precompute on the caller's stream first.
  const char* ov = getenv("ROCBTA_SI_OVERLAP");
  const hipStream_t ps = (ov && strcmp(ov, "0") == 0) ? s : side->s;
  RB_CHECK(hipEventRecord(side->ev[0], s));
  RB_CHECK(hipStreamWaitEvent(ps, side->ev[0], 0));
  for (int k = 0; k < nchunks; ++k) {
    const int hi = npre - k * SI_CHUNK;             // exclusive
    const int lo = std::max(0, hi - SI_CHUNK);      // inclusive
    const int nblk = hi - lo;
    // Linv enters G, Gt and S only as a triangular operand
    int rc = block_trtri(D + lo * bb, b, bb, Linv, b, bb, bi, nblk, W, ps, true, false);
    if (rc) return rc;
    GemmList pre;
    const int ng = std::min<int>(hi, (int)n - 1) - lo;
    if (ng > 0)
      push_term(pre.add(G + lo * bb, b, bi, bi, 0.0, TRI_NONE, ng, bb),
                term(L + lo * bb, b, false, Linv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER, bb, bb));
    if (a > 0)
      push_term(pre.add(Gt + lo * ab, b, ai, bi, 0.0, TRI_NONE, nblk, ab),
                term(R + lo * ab, b, false, Linv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER, ab, bb));
    // S_i is symmetric: lower tiles computed, mirrored to the upper part
    GemmProblem& psym = pre.add(D + lo * bb, b, bi, bi, 0.0, TRI_LOWER, nblk, bb);
    psym.mirror = 1;
    push_term(psym, term(Linv, b, true, Linv, b, false, bi, 1.0, TRI_UPPER, TRI_LOWER, bb, bb));
    rc = run_gemms(pre, ps);
    if (rc) return rc;
    RB_CHECK(hipEventRecord(side->ev[k + 1], ps));
  }
  auto chunk_of = [&](int64_t blk) { return (int)((npre - 1 - blk) / SI_CHUNK); };
  int waited = -1;
  auto need = [&](int64_t blk) -> int {
    const int k = chunk_of(blk);
    while (waited < k) {
      ++waited;
      RB_CHECK(hipStreamWaitEvent(s, side->ev[waited + 1], 0));
    }
    return ROCBTA_OK;
  };
  int rc = ROCBTA_OK;
  if (invert_last_block) {  // pobtasi.py:114-139
    rc = need(n - 1);
    if (rc) return rc;
    double* Rl = R + (n - 1) * ab;
    double* Dl = D + (n - 1) * bb;
    if (a > 0) {
      rc = block_trtri(T, a, 0, Tinv, a, a * a, ai, 1, Wt, s);
      if (rc) return rc;
      GemmList xt;
      push_term(xt.add(T, a, ai, ai, 0.0), term(Tinv, a, true, Tinv, a, false, ai, 1.0,
                                                 TRI_UPPER, TRI_LOWER));
      rc = run_gemms(xt, s);
      if (rc) return rc;
      GemmList xr;
      push_term(xr.add(Rl, b, ai, bi, 0.0),
                term(T, a, false, Gt + (n - 1) * ab, b, false, ai, -1.0));
      rc = run_gemms(xr, s);
      if (rc) return rc;
      GemmList xd;
      GemmProblem& pd = xd.add(Dl, b, bi, bi, 1.0, TRI_LOWER);
      pd.mirror = 1;
      push_term(pd, term(Rl, b, true, Gt + (n - 1) * ab, b, false, ai, -1.0));
      rc = run_gemms(xd, s);
      if (rc) return rc;
    }
  }
  // (3) backward recurrence (pobtasi.py:141-172) with the precomputed G's.
  // The arrow row block X_{t,i} = -X_{t,i+1} G_i - X_T Gt_i depends on
  // neither X_{i+1,i} nor X_ii, so it rides in the launch of X_{i+1,i+1}
  // (whose upper tiles leave CUs idle) instead of in the launch of
  // X_{i+1,i}, which then is exactly one 64-tile per CU at b = 1024.
  auto arrow_row = [&](GemmList& gl, int64_t i) {
    GemmProblem& pr = gl.add(R + i * ab, b, ai, bi, 0.0);
    pr.ksplit = ksplit;
    pr.partial = part_r;
    pr.counters = cnt_r;
    push_term(pr, term(R + (i + 1) * ab, b, false, G + i * bb, b, false, bi, -1.0));
    push_term(pr, term(T, a, false, Gt + i * ab, b, false, ai, -1.0));
  };
  if (a > 0 && n >= 2) {
    rc = need(n - 2);
    if (rc) return rc;
    GemmList first;
    arrow_row(first, n - 2);
    rc = run_gemms(first, s);
    if (rc) return rc;
  }
  for (int64_t i = n - 2; i >= 0; --i) {
    rc = need(i);
    if (rc) return rc;
    GemmList off;
    GemmProblem& pl = off.add(L + i * bb, b, bi, bi, 0.0);
    pl.ksplit = ksplit;
    pl.partial = part;
    pl.counters = cnt;
    push_term(pl, term(D + (i + 1) * bb, b, false, G + i * bb, b, false, bi, -1.0));
    if (a > 0) push_term(pl, term(R + (i + 1) * ab, b, true, Gt + i * ab, b, false, ai, -1.0));
    rc = run_gemms(off, s);
    if (rc) return rc;
    // X_ii = S_i - X_{i+1,i}^T G_i - X_{t,i}^T Gt_i is symmetric: lower
    // tiles computed, mirrored to the upper part.
    GemmList dia;
    GemmProblem& pd = dia.add(D + i * bb, b, bi, bi, 1.0, TRI_LOWER);
    pd.mirror = 1;
    pd.ksplit = ksplit;
    pd.partial = part;
    pd.counters = cnt;
    push_term(pd, term(L + i * bb, b, true, G + i * bb, b, false, bi, -1.0));
    if (a > 0) push_term(pd, term(R + i * ab, b, true, Gt + i * ab, b, false, ai, -1.0));
    if (a > 0 && i >= 1) {
      rc = need(i - 1);
      if (rc) return rc;
      arrow_row(dia, i - 1);
    }
    rc = run_gemms(dia, s);
    if (rc) return rc;
  }
  // every chunk has been waited for by now (the recurrence reaches block 0)
  while (waited + 1 < nchunks) {
    ++waited;
    RB_CHECK(hipStreamWaitEvent(s, side->ev[waited + 1], 0));
  }
  return ROCBTA_OK;
}

}  // extern "C"

namespace {

int trsv_tile(const double* Lt, int64_t ld, int m, double* B, int64_t ldb, int r, int trans,
              hipStream_t s) {
  if (m <= 0 || r <= 0) return ROCBTA_OK;
  trsv_tile_kernel<<<cdiv(r, SV_COLS), SV_THREADS, 0, s>>>(Lt, ld, m, B, ldb, r, trans);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

// Forward solve of one diagonal block (m x m, stride ld) on Bi (m x r), with
// the right-looking updates of `below` panels: {matrix (rows x m), target B}.
struct Below {
  const double* M;
  int64_t ld;
  int rows;
  double* Bt;
};

int forward_block(const double* Dm, int64_t ld, int m, double* Bi, int64_t ldb, int r,
                  const std::vector<Below>& below, hipStream_t s) {
  for (int k0 = 0; k0 < m; k0 += TT) {
    const int tk = std::min(TT, m - k0);
    int rc = trsv_tile(Dm + (int64_t)k0 * (ld + 1), ld, tk, Bi + (int64_t)k0 * ldb, ldb, r, 0, s);
    if (rc) return rc;
    GemmList upd;
    const double* yk = Bi + (int64_t)k0 * ldb;
    const int rem = m - k0 - tk;
    if (rem > 0)
      push_term(upd.add(Bi + (int64_t)(k0 + tk) * ldb, ldb, rem, r, 1.0),
                term(Dm + (int64_t)(k0 + tk) * ld + k0, ld, false, yk, ldb, false, tk, -1.0));
    for (const Below& bl : below)
      if (bl.rows > 0)
        push_term(upd.add(bl.Bt, ldb, bl.rows, r, 1.0),
                  term(bl.M + k0, bl.ld, false, yk, ldb, false, tk, -1.0));
    rc = run_gemms(upd, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

int backward_block(const double* Dm, int64_t ld, int m, double* Bi, int64_t ldb, int r,
                   hipStream_t s) {
  const int nt = cdiv(m, TT);
  for (int t = nt - 1; t >= 0; --t) {
    const int k0 = t * TT, tk = std::min(TT, m - k0);
    int rc = trsv_tile(Dm + (int64_t)k0 * (ld + 1), ld, tk, Bi + (int64_t)k0 * ldb, ldb, r, 1, s);
    if (rc) return rc;
    if (k0 == 0) continue;
    GemmList upd;
    push_term(upd.add(Bi, ldb, k0, r, 1.0),
              term(Dm + (int64_t)k0 * ld, ld, true, Bi + (int64_t)k0 * ldb, ldb, false, tk, -1.0));
    rc = run_gemms(upd, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

}  // namespace

extern "C" {

int rocbta_dpobtas(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                   const double* R, const double* T, double* B, int64_t r, int64_t ldb,
                   char trans, int partial, void* workspace, size_t workspace_bytes,
                   void* stream) {
  if (bad_shape(n, b, a) || !D || !B || (n > 1 && !L) || (a > 0 && (!R || !T)) || r < 0 ||
      ldb < r)
    return ROCBTA_ERR_ARG;
  if (trans != 'N' && trans != 'T' && trans != 'C') return ROCBTA_ERR_ARG;
  if (r == 0) return ROCBTA_OK;
  hipStream_t s = (hipStream_t)stream;
  const int algo = pobtas_algo(b, a, r, partial);
  if (algo == SOLVE_BLOCK) {
    if (!workspace || workspace_bytes < bs_layout(n, b, a, r).total) return ROCBTA_ERR_WORKSPACE;
    return pobtas_block(n, b, a, D, L, R, T, B, r, ldb, trans == 'N' ? 0 : 1, partial,
                        (char*)workspace, s);
  }
  if (algo == SOLVE_WIDE) {
    if (!workspace || workspace_bytes < wd_layout(n, b, a).total) return ROCBTA_ERR_WORKSPACE;
    return pobtas_wide(n, b, a, D, L, R, T, B, r, ldb, trans == 'N' ? 0 : 1, (char*)workspace, s);
  }
  if (algo == SOLVE_SWEEP) {
    if (!workspace || workspace_bytes < sw_layout(n, b, a, r).total) return ROCBTA_ERR_WORKSPACE;
    return pobtas_sweep(n, b, a, D, L, R, T, B, r, ldb, trans == 'N' ? 0 : 1, (char*)workspace, s);
  }
  if (workspace && workspace_bytes >= sizeof(int)) {
    set_int_kernel<<<1, 64, 0, s>>>((int*)workspace, 0);  // status word: no time-out
    RB_CHECK(hipGetLastError());
  }
  const int bi = (int)b, ai = (int)a, ri = (int)r;
  const int64_t bb = b * b, ab = a * b;
  double* Bt = B + n * b * ldb;  // tip rows
  if (trans == 'N') {  // pobtas.py:80-126
    for (int64_t i = 0; i + 1 < n; ++i) {
      std::vector<Below> below = {{L + i * bb, b, bi, B + (i + 1) * b * ldb}};
      if (a > 0) below.push_back({R + i * ab, b, ai, Bt});
      int rc = forward_block(D + i * bb, b, bi, B + i * b * ldb, ldb, ri, below, s);
      if (rc) return rc;
    }
    if (!partial) {
      std::vector<Below> below;
      if (a > 0) below.push_back({R + (n - 1) * ab, b, ai, Bt});
      int rc = forward_block(D + (n - 1) * bb, b, bi, B + (n - 1) * b * ldb, ldb, ri, below, s);
      if (rc) return rc;
      if (a > 0) {
        rc = forward_block(T, a, ai, Bt, ldb, ri, {}, s);
        if (rc) return rc;
      }
    }
  } else {  // pobtas.py:127-159
    if (!partial) {
      if (a > 0) {
        int rc = backward_block(T, a, ai, Bt, ldb, ri, s);
        if (rc) return rc;
        GemmList upd;
        push_term(upd.add(B + (n - 1) * b * ldb, ldb, bi, ri, 1.0),
                  term(R + (n - 1) * ab, b, true, Bt, ldb, false, ai, -1.0));
        rc = run_gemms(upd, s);
        if (rc) return rc;
      }
      int rc = backward_block(D + (n - 1) * bb, b, bi, B + (n - 1) * b * ldb, ldb, ri, s);
      if (rc) return rc;
    }
    for (int64_t i = n - 2; i >= 0; --i) {
      GemmList upd;
      GemmProblem& p = upd.add(B + i * b * ldb, ldb, bi, ri, 1.0);
      push_term(p, term(L + i * bb, b, true, B + (i + 1) * b * ldb, ldb, false, bi, -1.0));
      if (a > 0) push_term(p, term(R + i * ab, b, true, Bt, ldb, false, ai, -1.0));
      int rc = run_gemms(upd, s);
      if (rc) return rc;
      rc = backward_block(D + i * bb, b, bi, B + i * b * ldb, ldb, ri, s);
      if (rc) return rc;
    }
  }
  return ROCBTA_OK;
}

int64_t rocbta_dag_selfcheck(int64_t n, int64_t b, int64_t a) {
  if (bad_shape(n, b, a) || a > CT) return -1;
  const AfLayout lay = af_layout(n, b, a);
  double* const base = reinterpret_cast<double*>(uintptr_t(1) << 40);
  int* const fbase = reinterpret_cast<int*>(uintptr_t(1) << 44);
  PobtafChain P = chain_params(n, b, a, base, base + ((int64_t)1 << 36), base + ((int64_t)2 << 36),
                               base + ((int64_t)3 << 36), nullptr, nullptr, lay);
  P.hflags = fbase;
  P.cflags = fbase + ((int64_t)1 << 32);
  int64_t bad = 0;
  auto in_range = [&](const CTile& t, const double* arr, int64_t count) {
    const int64_t off = t.p - arr;
    if (t.rows <= 0 || t.cols <= 0 || t.rows > CT || t.cols > CT) return false;
    return off >= 0 && off + (int64_t)(t.rows - 1) * t.ld + (t.cols - 1) < count;
  };
  auto tile_ok = [&](const CTile& t) {
    return in_range(t, P.D, n * b * b) || in_range(t, P.L, (n - 1) * b * b) ||
           in_range(t, P.R, n * a * b);
  };
  // a helper may wait on a helper task strictly earlier in the queue, or on a
  // chain output of a step before its stage
  auto wait_ok = [&](const int* f, int task) {
    const int stage = task / P.slots;
    if (f >= P.cflags) {
      const int64_t idx = f - P.cflags;
      return idx >= 0 && idx < 2 * (int64_t)P.NT && idx / 2 <= stage - 1;
    }
    const int64_t idx = f - P.hflags;
    return idx >= 0 && idx < task;
  };
  std::vector<int> seen_trsm, seen_other;
  int64_t nwork = 0;
  for (int t = 0; t < P.ntasks; ++t) {
    int r, j;
    const int kind = chain_decode(P, t, r, j);
    if (kind == HK_NOOP) continue;
    ++nwork;
    if (kind == HK_TIP) {
      if (t != t_tipc(P)) ++bad;
      for (int i = 0; i < P.n; ++i)
        if (!wait_ok(P.hflags + t_tippc(P, i), t)) ++bad;
      continue;
    }
    if (kind == HK_TIPP) {
      if (t != t_tippc(P, j)) ++bad;
      for (int c = 0; c < P.nt; ++c)
        if (!wait_ok(P.hflags + t_arr(P, j * P.nt + c), t)) ++bad;
      continue;
    }
    int back = -1;
    if (kind == HK_TRSM) back = t_trsm(P, r, j);
    if (kind == HK_ARR) back = t_arr(P, j);
    if (kind == HK_PX) back = t_px(P, j);
    if (kind == HK_PD) back = t_pd(P, j);
    if (back != t) ++bad;
    if (kind == HK_TRSM && !(r >= j + 2 && r < band_end(P, j))) ++bad;
    const ContribRange cr = chain_contribs(P, kind, r, j);
    for (int k = cr.k0; k < cr.k1; ++k) {
      const CTile ta = cr.ra < 0 ? tile_arrow(P, k) : tile_g(P, cr.ra, k);
      if (!tile_ok(ta) || !tile_ok(tile_g(P, cr.jb, k))) ++bad;
      if (!wait_ok(factor_flag(P, cr.ra, k), t) || !wait_ok(factor_flag(P, cr.jb, k), t)) ++bad;
      if (cr.ra >= 0 && !(cr.ra > k)) ++bad;
      if (!(cr.jb > k)) ++bad;
    }
    if ((kind == HK_TRSM || kind == HK_ARR) && !wait_ok(P.cflags + 2 * j, t)) ++bad;
    CTile out = kind == HK_TRSM ? tile_g(P, r, j)
                : kind == HK_ARR ? tile_arrow(P, j)
                : kind == HK_PX  ? tile_g(P, j + 1, j)
                                 : tile_g(P, j, j);
    if (!tile_ok(out)) ++bad;
    if (kind == HK_TRSM && r / P.nt == j / P.nt && !tile_ok(tile_mirror(P, r, j))) ++bad;
  }
  // chain: step g waits PD(g) (g >= 1) and PX(g) when a sub-diagonal exists;
  // each must be a decoded task of the right kind, and it may only depend on
  // chain steps < g, which wait_ok enforces through its stage.
  for (int g = 0; g < P.NT; ++g) {
    int r, j;
    if (g > 0) {
      const int tpd = t_pd(P, g);
      if (chain_decode(P, tpd, r, j) != HK_PD || j != g || tpd / P.slots > g) ++bad;
    }
    if (g + 1 < band_end(P, g)) {
      const int tpx = t_px(P, g);
      if (chain_decode(P, tpx, r, j) != HK_PX || j != g || tpx / P.slots > g) ++bad;
      if (!tile_ok(tile_g(P, g + 1, g))) ++bad;
    }
    if (!tile_ok(tile_g(P, g, g))) ++bad;
  }
  // every lower band tile is produced exactly once (chain or helper)
  int64_t expect = 0;
  for (int j = 0; j < P.NT; ++j) expect += std::max(0, band_end(P, j) - (j + 2)) + (P.ha ? 1 : 0);
  int64_t got = 0;
  for (int t = 0; t < P.ntasks; ++t) {
    int r, j;
    const int kind = chain_decode(P, t, r, j);
    if (kind == HK_TRSM || kind == HK_ARR) ++got;
  }
  if (got != expect) ++bad;
  (void)nwork;
  return bad;
}

int rocbta_mpobtaf(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                   int factorize_last_block, int* info, void* workspace, size_t workspace_bytes,
                   void* stream) {
  MixedScope mixed(true);
  return rocbta_dpobtaf(n, b, a, D, L, R, T, factorize_last_block, info, workspace,
                        workspace_bytes, stream);
}

int rocbta_mpobtasi(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                    int invert_last_block, void* workspace, size_t workspace_bytes, void* stream) {
  MixedScope mixed(true);
  return rocbta_dpobtasi(n, b, a, D, L, R, T, invert_last_block, workspace, workspace_bytes,
                         stream);
}

int rocbta_dresidual(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                     const double* R, const double* T, const double* X, const double* B,
                     double* Out, int64_t r, int64_t ld, void* workspace, size_t workspace_bytes,
                     void* stream) {
  if (bad_shape(n, b, a) || !D || !X || !B || !Out || (n > 1 && !L) || (a > 0 && (!R || !T)) ||
      r < 1 || ld < r || Out == X)
    return ROCBTA_ERR_ARG;
  if (workspace_bytes < rocbta_workspace_bytes(ROCBTA_OP_RESIDUAL, n, b, a, r) ||
      (a > 0 && !workspace))
    return ROCBTA_ERR_WORKSPACE;
  ResidP P{};
  P.D = D;
  P.L = L;
  P.R = R;
  P.T = T;
  P.X = X;
  P.B = B;
  P.Out = Out;
  P.ld = ld;
  P.n = (int)n;
  P.b = (int)b;
  P.a = (int)a;
  P.nt = cdiv(b, SW_T);
  P.ta = cdiv(a, SW_T);
  P.NT = (int)(n * P.nt);
  P.r = (int)r;
  P.part = (double*)workspace;
  const int rc = sweep_rc(r);
  const int nch = cdiv(r, rc);
  hipStream_t s = (hipStream_t)stream;
  const dim3 g1((unsigned)(P.NT + (a > 0 ? n : 0)), (unsigned)nch), g2((unsigned)P.ta, (unsigned)nch);
  switch (rc) {
    case 1:
      residual_rows_kernel<1><<<g1, SW_THREADS, 0, s>>>(P);
      if (a > 0) residual_tip_kernel<1><<<g2, SW_THREADS, 0, s>>>(P);
      break;
    case 4:
      residual_rows_kernel<4><<<g1, SW_THREADS, 0, s>>>(P);
      if (a > 0) residual_tip_kernel<4><<<g2, SW_THREADS, 0, s>>>(P);
      break;
    default:
      residual_rows_kernel<8><<<g1, SW_THREADS, 0, s>>>(P);
      if (a > 0) residual_tip_kernel<8><<<g2, SW_THREADS, 0, s>>>(P);
      break;
  }
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

int rocbta_dgrad(int64_t n, int64_t b, int64_t a, const double* XD, const double* XL,
                 const double* XR, const double* XT, const double* x, int64_t r, int64_t ldx,
                 const double* gl, const double* gq, double* GD, double* GL, double* GR,
                 double* GT, double* gy, void* stream) {
  if (bad_shape(n, b, a) || !XD || !GD || (n > 1 && (!XL || !GL)) ||
      (a > 0 && (!XR || !XT || !GR || !GT)) || r < 0 || (r > 0 && (!x || ldx < r)))
    return ROCBTA_ERR_ARG;
  if (gq && r == 0) return ROCBTA_ERR_ARG;
  GradP P{};
  P.XD = XD;
  P.XL = XL;
  P.XR = XR;
  P.XT = XT;
  P.x = x;
  P.ldx = ldx;
  P.r = (int)r;
  P.n = n;
  P.b = b;
  P.a = a;
  P.gl = gl;
  P.gq = gq;
  P.GD = GD;
  P.GL = GL;
  P.GR = GR;
  P.GT = GT;
  P.gy = gy;
  P.nD = n * b * b;
  P.nL = (n - 1) * b * b;
  P.nR = n * a * b;
  P.nT = a * a;
  P.nY = gy ? (n * b + a) * r : 0;
  const int64_t total = P.nD + P.nL + P.nR + P.nT + P.nY;
  hipStream_t s = (hipStream_t)stream;
  const int grid = (int)std::min<int64_t>(8 * (int64_t)device_cus(), (total + 255) / 256);
  if (grid > 0) grad_assemble_kernel<<<grid, 256, 0, s>>>(P);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

int rocbta_dquad(int64_t rows, int64_t r, const double* y, const double* x, int64_t ld,
                 double* out, void* stream) {
  if (rows < 0 || r < 1 || ld < r || !y || !x || !out) return ROCBTA_ERR_ARG;
  quad_kernel<<<1, 256, 0, (hipStream_t)stream>>>(y, x, rows, (int)r, ld, out);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

int rocbta_dlogdet(int64_t n, int64_t b, int64_t a, const double* D, const double* T,
                   double* out, void* workspace, size_t workspace_bytes, void* stream) {
  if (bad_shape(n, b, a) || !D || !out || (a > 0 && !T)) return ROCBTA_ERR_ARG;
  if (!workspace || workspace_bytes < rocbta_workspace_bytes(ROCBTA_OP_LOGDET, n, b, a, 0))
    return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const int nparts = logdet_parts(n, b, a);
  double* part = (double*)workspace;
  logdet_part_kernel<<<nparts, 256, 0, s>>>(D, n, b, T, a, part);
  RB_CHECK(hipGetLastError());
  logdet_finish_kernel<<<1, 256, 0, s>>>(part, nparts, out);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

// -------------------------------------------------------------- permuted
// The "middle partition" forms of the distributed algorithm: the first and
// last diagonal blocks and the tip are left to a reduced system, and a top
// buffer column (buffer[i] = the coupling of block i with block 0) carries
// the fill-in.  Block recurrences over the same primitives as the natural
// forms: the blocked diagonal factorization with its panel solves
// (factor_block_column), the batched block inverse (block_trtri), grouped
// GEMMs and the 64-row tile solves.  a = 0 gives the BT forms.

int rocbta_dpobtaf_permuted(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                            double* T, double* buffer, int* info, void* workspace,
                            size_t workspace_bytes, void* stream) {
  if (bad_shape(n, b, a) || !D || !buffer || !info || (n > 1 && !L) || (a > 0 && (!R || !T)))
    return ROCBTA_ERR_ARG;
  if (!workspace || workspace_bytes < rocbta_workspace_bytes(ROCBTA_OP_POBTAF_PERMUTED, n, b, a, 0))
    return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  set_int_kernel<<<1, 64, 0, s>>>(info, 0);
  RB_CHECK(hipGetLastError());
  double* tinv = (double*)workspace;
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  // buffer[1] = A_{1,0}^T, the top block's coupling with block 1 (pobtaf.py:204)
  if (n >= 2) {
    int rc = transpose_block(L, buffer + bb, b, s);
    if (rc) return rc;
  }
  for (int64_t i = 1; i + 1 < n; ++i) {  // pobtaf.py:207-281
    // L_ii = chol(A_ii); L_{i+1,i}, L_{top,i} = buffer[i], L_{t,i} = R_i
    // times L_ii^{-T}
    std::vector<Panel> panels = {{L + i * bb, bi, b}, {buffer + i * bb, bi, b}};
    if (a > 0) panels.push_back({R + i * ab, ai, b});
    int rc = factor_block_column(D + i * bb, b, bi, panels, i * b, info, tinv, s);
    if (rc) return rc;
    const double* Li = L + i * bb;
    const double* Bi = buffer + i * bb;
    const double* Ri = R + i * ab;
    // Schur updates (independent outputs in one group).  Block i+1 is
    // factored next unless it is the last one, which keeps the full
    // symmetric update, like the top block D_0 and the tip.
    GemmList upd;
    const bool last = (i + 2 == n);
    push_term(upd.add(D + (i + 1) * bb, b, bi, bi, 1.0, last ? TRI_NONE : TRI_LOWER),
              term(Li, b, false, Li, b, true, bi, -1.0));
    push_term(upd.add(D, b, bi, bi, 1.0), term(Bi, b, false, Bi, b, true, bi, -1.0));
    push_term(upd.add(buffer + (i + 1) * bb, b, bi, bi, 0.0), term(Bi, b, false, Li, b, true, bi, -1.0));
    if (a > 0) {
      push_term(upd.add(R + (i + 1) * ab, b, ai, bi, 1.0), term(Ri, b, false, Li, b, true, bi, -1.0));
      push_term(upd.add(T, a, ai, ai, 1.0), term(Ri, b, false, Ri, b, true, bi, -1.0));
      push_term(upd.add(R, b, ai, bi, 1.0), term(Ri, b, false, Bi, b, true, bi, -1.0));
    }
    rc = run_gemms(upd, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

int rocbta_dpobtasi_permuted(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                             double* T, double* buffer, void* workspace, size_t workspace_bytes,
                             void* stream) {
  if (bad_shape(n, b, a) || !D || !buffer || (n > 1 && !L) || (a > 0 && (!R || !T)))
    return ROCBTA_ERR_ARG;
  const PsiLayout lay = psi_layout(n, b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double* Linv = (double*)(ws + lay.linv);
  double* Lt = (double*)(ws + lay.lt);
  double* Bt = (double*)(ws + lay.bt);
  double* Rt = (double*)(ws + lay.rt);
  double* WL = (double*)(ws + lay.wl);
  double* WB = (double*)(ws + lay.wb);
  double* WR = (double*)(ws + lay.wr);
  double* WD = (double*)(ws + lay.wd);
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  if (n >= 3) {
    // L_ii^{-1} of the middle blocks (pobtasi.py:205-209), batched
    int rc = block_trtri(D + bb, b, bb, Linv + bb, b, bb, bi, (int)(n - 2), (double*)(ws + lay.w), s,
                         true);
    if (rc) return rc;
  }
  for (int64_t i = n - 2; i >= 1; --i) {  // pobtasi.py:199-240
    const double* Li_inv = Linv + i * bb;
    // the factor blocks this step overwrites but still reads
    RB_CHECK(hipMemcpyAsync(Lt, L + i * bb, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    RB_CHECK(hipMemcpyAsync(Bt, buffer + i * bb, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (a > 0)
      RB_CHECK(hipMemcpyAsync(Rt, R + i * ab, ab * sizeof(double), hipMemcpyDeviceToDevice, s));
    const double* X1 = D + (i + 1) * bb;      // X_{i+1,i+1}
    const double* B1 = buffer + (i + 1) * bb;  // X_{top,i+1}
    const double* R1 = R + (i + 1) * ab;       // X_{t,i+1}
    // bracketed sums of X_{i+1,i}, X_{top,i}, X_{t,i}
    GemmList w1;
    push_term(w1.add(WL, b, bi, bi, 0.0), term(B1, b, true, Bt, b, false, bi, -1.0));
    push_term(w1.problems_back(), term(X1, b, false, Lt, b, false, bi, -1.0));
    push_term(w1.add(WB, b, bi, bi, 0.0), term(B1, b, false, Lt, b, false, bi, -1.0));
    push_term(w1.problems_back(), term(D, b, false, Bt, b, false, bi, -1.0));
    if (a > 0) {
      push_term(w1.add(WR, b, ai, bi, 0.0), term(R1, b, false, Lt, b, false, bi, -1.0));
      push_term(w1.problems_back(), term(R, b, false, Bt, b, false, bi, -1.0));
    }
    int rc = run_gemms(w1, s);
    if (rc) return rc;
    if (a > 0) {
      GemmList w2;
      push_term(w2.add(WL, b, bi, bi, 1.0), term(R1, b, true, Rt, b, false, ai, -1.0));
      push_term(w2.add(WB, b, bi, bi, 1.0), term(R, b, true, Rt, b, false, ai, -1.0));
      push_term(w2.add(WR, b, ai, bi, 1.0), term(T, a, false, Rt, b, false, ai, -1.0));
      rc = run_gemms(w2, s);
      if (rc) return rc;
    }
    // times L_ii^{-1} (lower triangular: K trimmed)
    GemmList w3;
    push_term(w3.add(L + i * bb, b, bi, bi, 0.0),
              term(WL, b, false, Li_inv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER));
    push_term(w3.add(buffer + i * bb, b, bi, bi, 0.0),
              term(WB, b, false, Li_inv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER));
    if (a > 0)
      push_term(w3.add(R + i * ab, b, ai, bi, 0.0),
                term(WR, b, false, Li_inv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER));
    rc = run_gemms(w3, s);
    if (rc) return rc;
    // X_ii = (L_ii^{-T} - X_{i+1,i}^T L_{i+1,i} - X_{top,i}^T L_{top,i} - X_{t,i}^T L_{t,i}) L_ii^{-1}
    GemmList w4;
    push_term(w4.add(WD, b, bi, bi, 0.0), term(L + i * bb, b, true, Lt, b, false, bi, -1.0));
    push_term(w4.problems_back(), term(buffer + i * bb, b, true, Bt, b, false, bi, -1.0));
    rc = run_gemms(w4, s);
    if (rc) return rc;
    if (a > 0) {
      GemmList w5;
      push_term(w5.add(WD, b, bi, bi, 1.0), term(R + i * ab, b, true, Rt, b, false, ai, -1.0));
      rc = run_gemms(w5, s);
      if (rc) return rc;
    }
    GemmList w6;
    push_term(w6.add(D + i * bb, b, bi, bi, 0.0),
              term(WD, b, false, Li_inv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER));
    push_term(w6.problems_back(),
              term(Li_inv, b, true, Li_inv, b, false, bi, 1.0, TRI_UPPER, TRI_LOWER));
    rc = run_gemms(w6, s);
    if (rc) return rc;
  }
  // X_{1,0} = X_{top,1}^T back to the tridiagonal storage (pobtasi.py:244)
  if (n >= 2) return transpose_block(buffer + bb, L, b, s);
  return ROCBTA_OK;
}

int rocbta_dpobtas_permuted(int64_t n, int64_t b, int64_t a, const double* D, const double* L,
                            const double* R, const double* T, const double* buffer, double* B,
                            int64_t r, int64_t ldb, char trans, void* workspace,
                            size_t workspace_bytes, void* stream) {
  (void)T;
  if (bad_shape(n, b, a) || !D || !B || !buffer || (n > 1 && !L) || (a > 0 && !R) || r < 0 ||
      ldb < r)
    return ROCBTA_ERR_ARG;
  if (trans != 'N' && trans != 'T' && trans != 'C') return ROCBTA_ERR_ARG;
  if (r == 0) return ROCBTA_OK;
  hipStream_t s = (hipStream_t)stream;
  if (workspace && workspace_bytes >= sizeof(int)) {
    set_int_kernel<<<1, 64, 0, s>>>((int*)workspace, 0);  // status word: no time-out
    RB_CHECK(hipGetLastError());
  }
  const int bi = (int)b, ai = (int)a, ri = (int)r;
  const int64_t bb = b * b, ab = a * b;
  double* Bt = B + n * b * ldb;  // tip rows
  if (trans == 'N') {  // pobtas.py:179-204
    for (int64_t i = 1; i + 1 < n; ++i) {
      std::vector<Below> below = {{L + i * bb, b, bi, B + (i + 1) * b * ldb},
                                  {buffer + i * bb, b, bi, B}};
      if (a > 0) below.push_back({R + i * ab, b, ai, Bt});
      int rc = forward_block(D + i * bb, b, bi, B + i * b * ldb, ldb, ri, below, s);
      if (rc) return rc;
    }
    return ROCBTA_OK;
  }
  for (int64_t i = n - 2; i >= 1; --i) {  // pobtas.py:205-216
    double* Bi = B + i * b * ldb;
    GemmList upd;
    GemmProblem& p = upd.add(Bi, ldb, bi, ri, 1.0);
    push_term(p, term(L + i * bb, b, true, B + (i + 1) * b * ldb, ldb, false, bi, -1.0));
    push_term(p, term(buffer + i * bb, b, true, B, ldb, false, bi, -1.0));
    int rc = run_gemms(upd, s);
    if (rc) return rc;
    if (a > 0) {
      GemmList upd2;
      push_term(upd2.add(Bi, ldb, bi, ri, 1.0), term(R + i * ab, b, true, Bt, ldb, false, ai, -1.0));
      rc = run_gemms(upd2, s);
      if (rc) return rc;
    }
    rc = backward_block(D + i * bb, b, bi, Bi, ldb, ri, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

// ----------------------------------------------------- non-SPD (ddbt*)
// Block LU Schur-complement sweep and selected inversion of diagonally
// dominant BTA / BT matrices (no symmetry): blocked Gauss-Jordan inverses
// of the diagonal blocks and grouped GEMMs of the updates.

int rocbta_dddbtasc(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                    double* Rl, double* Ru, double* T, int invert_last_block, int direction,
                    void* workspace, size_t workspace_bytes, void* stream) {
  if (bad_shape(n, b, a) || !D || (n > 1 && (!Lo || !Up)) || (a > 0 && (!Rl || !Ru || !T)))
    return ROCBTA_ERR_ARG;
  if (direction != 0 && (direction != 1 || a > 0)) return ROCBTA_ERR_ARG;  // upward: BT only
  const DdLayout lay = dd_layout(b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double* W = (double*)(ws + lay.w);
  double* V = (double*)(ws + lay.v);
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  int rc = ROCBTA_OK;
  if (direction == 1) {  // ddbtsc.py:202-221
    for (int64_t i = n - 1; i >= 1; --i) {
      rc = block_inverse(D + i * bb, b, bi, ws, lay, s);
      if (rc) return rc;
      GemmList g1;
      push_term(g1.add(W, b, bi, bi, 0.0), term(D + i * bb, b, false, Lo + (i - 1) * bb, b, false, bi, 1.0));
      rc = run_gemms(g1, s);
      if (rc) return rc;
      GemmList g2;
      push_term(g2.add(D + (i - 1) * bb, b, bi, bi, 1.0),
                term(Up + (i - 1) * bb, b, false, W, b, false, bi, -1.0));
      rc = run_gemms(g2, s);
      if (rc) return rc;
    }
    return invert_last_block ? block_inverse(D, b, bi, ws, lay, s) : ROCBTA_OK;
  }
  for (int64_t i = 1; i < n; ++i) {  // ddbtasc.py:199-233, ddbtsc.py:188-195
    double* Dp = D + (i - 1) * bb;
    rc = block_inverse(Dp, b, bi, ws, lay, s);
    if (rc) return rc;
    // W = A_{i-1,i-1}^{-1} A_{i-1,i}, V = A_{i-1,i-1}^{-1} A_{i-1,t}
    GemmList g1;
    push_term(g1.add(W, b, bi, bi, 0.0), term(Dp, b, false, Up + (i - 1) * bb, b, false, bi, 1.0));
    if (a > 0)
      push_term(g1.add(V, ai, bi, ai, 0.0), term(Dp, b, false, Ru + (i - 1) * ab, ai, false, bi, 1.0));
    rc = run_gemms(g1, s);
    if (rc) return rc;
    GemmList g2;
    push_term(g2.add(D + i * bb, b, bi, bi, 1.0), term(Lo + (i - 1) * bb, b, false, W, b, false, bi, -1.0));
    if (a > 0) {
      push_term(g2.add(Rl + i * ab, b, ai, bi, 1.0), term(Rl + (i - 1) * ab, b, false, W, b, false, bi, -1.0));
      push_term(g2.add(Ru + i * ab, ai, bi, ai, 1.0), term(Lo + (i - 1) * bb, b, false, V, ai, false, bi, -1.0));
      push_term(g2.add(T, a, ai, ai, 1.0), term(Rl + (i - 1) * ab, b, false, V, ai, false, bi, -1.0));
    }
    rc = run_gemms(g2, s);
    if (rc) return rc;
  }
  if (!invert_last_block) return ROCBTA_OK;
  double* Dl = D + (n - 1) * bb;  // ddbtasc.py:235-241
  rc = block_inverse(Dl, b, bi, ws, lay, s);
  if (rc || a == 0) return rc;
  GemmList g1;
  push_term(g1.add(V, ai, bi, ai, 0.0), term(Dl, b, false, Ru + (n - 1) * ab, ai, false, bi, 1.0));
  rc = run_gemms(g1, s);
  if (rc) return rc;
  GemmList g2;
  push_term(g2.add(T, a, ai, ai, 1.0), term(Rl + (n - 1) * ab, b, false, V, ai, false, bi, -1.0));
  rc = run_gemms(g2, s);
  if (rc) return rc;
  return block_inverse(T, a, ai, ws, lay, s);
}

int rocbta_dddbtasci(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                     double* Rl, double* Ru, double* T, int invert_last_block, int direction,
                     void* workspace, size_t workspace_bytes, void* stream) {
  if (bad_shape(n, b, a) || !D || (n > 1 && (!Lo || !Up)) || (a > 0 && (!Rl || !Ru || !T)))
    return ROCBTA_ERR_ARG;
  if (direction != 0 && (direction != 1 || a > 0)) return ROCBTA_ERR_ARG;
  const DdLayout lay = dd_layout(b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double* B1 = (double*)(ws + lay.b1);
  double* B2 = (double*)(ws + lay.b2);
  double* C1 = (double*)(ws + lay.c1);
  double* C2 = (double*)(ws + lay.c2);
  double* E = (double*)(ws + lay.e);
  double* Gm = (double*)(ws + lay.g);
  double* Dc = (double*)(ws + lay.dc);
  double* Y = (double*)(ws + lay.y);
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  int rc = ROCBTA_OK;
  if (direction == 1) {  // ddbtsci.py:208-237
    for (int64_t i = 1; i < n; ++i) {
      double* Di = D + i * bb;
      double* Dm = D + (i - 1) * bb;
      double* Lm = Lo + (i - 1) * bb;
      double* Um = Up + (i - 1) * bb;
      RB_CHECK(hipMemcpyAsync(Dc, Di, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
      // B1 = A_{i,i-1} X_{i-1,i-1}; C1 = X_{i-1,i-1} A_{i-1,i}
      GemmList g1;
      push_term(g1.add(B1, b, bi, bi, 0.0), term(Lm, b, false, Dm, b, false, bi, 1.0));
      push_term(g1.add(C1, b, bi, bi, 0.0), term(Dm, b, false, Um, b, false, bi, 1.0));
      rc = run_gemms(g1, s);
      if (rc) return rc;
      // E = B1 A_{i-1,i} (old), then the new off-diagonal blocks
      GemmList g2;
      push_term(g2.add(E, b, bi, bi, 0.0), term(B1, b, false, Um, b, false, bi, 1.0));
      push_term(g2.add(Lm, b, bi, bi, 0.0), term(Dc, b, false, B1, b, false, bi, -1.0));
      rc = run_gemms(g2, s);
      if (rc) return rc;
      GemmList g3;
      push_term(g3.add(Um, b, bi, bi, 0.0), term(C1, b, false, Dc, b, false, bi, -1.0));
      push_term(g3.add(Gm, b, bi, bi, 0.0), term(Dc, b, false, E, b, false, bi, 1.0));
      rc = run_gemms(g3, s);
      if (rc) return rc;
      GemmList g4;
      push_term(g4.add(Di, b, bi, bi, 1.0), term(Gm, b, false, Dc, b, false, bi, 1.0));
      rc = run_gemms(g4, s);
      if (rc) return rc;
    }
    return ROCBTA_OK;
  }
  if (invert_last_block && a > 0) {  // ddbtasci.py:207-219
    double* Dl = D + (n - 1) * bb;
    double* Rll = Rl + (n - 1) * ab;
    double* Rul = Ru + (n - 1) * ab;
    GemmList g1;
    push_term(g1.add(B2, ai, bi, ai, 0.0), term(Dl, b, false, Rul, ai, false, bi, 1.0));
    push_term(g1.add(Y, b, ai, bi, 0.0), term(Rll, b, false, Dl, b, false, bi, 1.0));
    rc = run_gemms(g1, s);
    if (rc) return rc;
    GemmList g2;
    push_term(g2.add(Rll, b, ai, bi, 0.0), term(T, a, false, Y, b, false, ai, -1.0));
    push_term(g2.add(Rul, ai, bi, ai, 0.0), term(B2, ai, false, T, a, false, ai, -1.0));
    rc = run_gemms(g2, s);
    if (rc) return rc;
    GemmList g3;
    push_term(g3.add(Dl, b, bi, bi, 1.0), term(B2, ai, false, Rll, b, false, ai, -1.0));
    rc = run_gemms(g3, s);
    if (rc) return rc;
  }
  for (int64_t i = n - 2; i >= 0; --i) {  // ddbtasci.py:221-256, ddbtsci.py:187-205
    double* Di = D + i * bb;
    double* D1 = D + (i + 1) * bb;
    double* Li = Lo + i * bb;
    double* Ui = Up + i * bb;
    double* Rli = Rl + i * ab;
    double* Rui = Ru + i * ab;
    double* Rl1 = Rl + (i + 1) * ab;
    double* Ru1 = Ru + (i + 1) * ab;
    RB_CHECK(hipMemcpyAsync(Dc, Di, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    GemmList g1;
    GemmProblem& pb1 = g1.add(B1, b, bi, bi, 0.0);
    push_term(pb1, term(Ui, b, false, D1, b, false, bi, 1.0));
    GemmProblem& pc1 = g1.add(C1, b, bi, bi, 0.0);
    push_term(pc1, term(D1, b, false, Li, b, false, bi, 1.0));
    if (a > 0) {
      push_term(pb1, term(Rui, ai, false, Rl1, b, false, ai, 1.0));
      push_term(pc1, term(Ru1, ai, false, Rli, b, false, ai, 1.0));
      GemmProblem& pb2 = g1.add(B2, ai, bi, ai, 0.0);
      push_term(pb2, term(Ui, b, false, Ru1, ai, false, bi, 1.0));
      push_term(pb2, term(Rui, ai, false, T, a, false, ai, 1.0));
      GemmProblem& pc2 = g1.add(C2, b, ai, bi, 0.0);
      push_term(pc2, term(Rl1, b, false, Li, b, false, bi, 1.0));
      push_term(pc2, term(T, a, false, Rli, b, false, ai, 1.0));
    }
    rc = run_gemms(g1, s);
    if (rc) return rc;
    // E = B1 A_{i+1,i} + B2 A_{t,i} (old blocks); X_{i,i+1}, X_{i,t}
    GemmList g2;
    GemmProblem& pe = g2.add(E, b, bi, bi, 0.0);
    push_term(pe, term(B1, b, false, Li, b, false, bi, 1.0));
    push_term(g2.add(Ui, b, bi, bi, 0.0), term(Dc, b, false, B1, b, false, bi, -1.0));
    if (a > 0) {
      push_term(pe, term(B2, ai, false, Rli, b, false, ai, 1.0));
      push_term(g2.add(Rui, ai, bi, ai, 0.0), term(Dc, b, false, B2, ai, false, bi, -1.0));
    }
    rc = run_gemms(g2, s);
    if (rc) return rc;
    // X_{i+1,i}, X_{t,i}, and Gm = X_ii E
    GemmList g3;
    push_term(g3.add(Li, b, bi, bi, 0.0), term(C1, b, false, Dc, b, false, bi, -1.0));
    if (a > 0) push_term(g3.add(Rli, b, ai, bi, 0.0), term(C2, b, false, Dc, b, false, bi, -1.0));
    push_term(g3.add(Gm, b, bi, bi, 0.0), term(Dc, b, false, E, b, false, bi, 1.0));
    rc = run_gemms(g3, s);
    if (rc) return rc;
    // X_ii = D_i + D_i E D_i
    GemmList g4;
    push_term(g4.add(Di, b, bi, bi, 1.0), term(Gm, b, false, Dc, b, false, bi, 1.0));
    rc = run_gemms(g4, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

// A X A^T = B forms: the same sweeps with B's blocks carried along.
namespace {
struct Gq {  // one grouped launch, built problem by problem
  GemmList l;
  GemmProblem* cur = nullptr;
  Gq& p(double* C, int64_t ldc, int64_t M, int64_t N, double beta) {
    cur = &l.add(C, ldc, (int)M, (int)N, beta);
    return *this;
  }
  // alpha * op(A) op(B), K = k
  Gq& t(const double* A, int64_t lda, bool tA, const double* B, int64_t ldb, bool tB, int64_t k,
        double alpha) {
    push_term(*cur, term(A, lda, tA, B, ldb, tB, (int)k, alpha));
    return *this;
  }
  int run(hipStream_t s) { return run_gemms(l, s); }
};
}  // namespace

// Block views with a signed stride: the upward BT forms run the downward
// recurrences on the block-reversed matrix J A J (J reverses the block
// order), whose diagonal blocks are A's in reverse order and whose lower
// (upper) blocks are A's upper (lower) blocks in reverse order; B likewise.
// The results are the same blocks of A^{-1} and A^{-1} B A^{-T}.
struct Bv {
  double* p;
  int64_t s;
  double* operator()(int64_t i) const { return p + i * s; }
};

static Bv fwd_view(double* p, int64_t bb) { return {p, bb}; }
static Bv rev_view(double* p, int64_t count, int64_t bb) {
  return count > 0 ? Bv{p + (count - 1) * bb, -bb} : Bv{p, bb};
}

static int dd_quadratic_sc(int64_t n, int64_t b, int64_t a, Bv Dv, Bv Lov, Bv Upv,
                           double* Rl, double* Ru, double* T, Bv BDv, Bv BLov, Bv BUpv,
                           double* BRl, double* BRu, double* BT, int invert_last_block,
                           void* workspace, size_t workspace_bytes, void* stream) {
  const DdLayout lay = dd_layout(b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  auto P = [&](size_t off) { return (double*)(ws + off); };
  double *W = P(lay.w), *V = P(lay.v), *Z = P(lay.z), *P1 = P(lay.p1), *P2 = P(lay.p2);
  double *M1 = P(lay.m1), *M2 = P(lay.m2), *Y1 = P(lay.bbt);
  const int64_t ab = a * b;
  int rc = ROCBTA_OK;
  for (int64_t i = 1; i < n; ++i) {  // ddbtasc.py:351-445
    double* Di = Dv(i - 1);
    const double* L = Lov(i - 1);
    const double* U = Upv(i - 1);
    const double* R = Rl + (i - 1) * ab;
    const double* Rur = Ru + (i - 1) * ab;
    rc = block_inverse(Di, b, (int)b, ws, lay, s);
    if (rc) return rc;
    Gq g1;
    g1.p(W, b, b, b, 0.0).t(Di, b, false, U, b, false, b, 1.0);
    g1.p(Z, b, b, b, 0.0).t(Di, b, false, BDv(i - 1), b, false, b, 1.0);
    g1.p(P1, b, b, b, 0.0).t(L, b, false, Di, b, false, b, 1.0);
    if (a > 0) {
      g1.p(V, a, b, a, 0.0).t(Di, b, false, Rur, a, false, b, 1.0);
      g1.p(P2, b, a, b, 0.0).t(R, b, false, Di, b, false, b, 1.0);
    }
    if ((rc = g1.run(s))) return rc;
    Gq g2;  // A's Schur updates; B_{i-1} <- A_{i-1}^{-1} B_{i-1} A_{i-1}^{-T}
    g2.p(Dv(i), b, b, b, 1.0).t(L, b, false, W, b, false, b, -1.0);
    g2.p(BDv(i - 1), b, b, b, 0.0).t(Z, b, false, Di, b, true, b, 1.0);
    if (a > 0) {
      g2.p(Rl + i * ab, b, a, b, 1.0).t(R, b, false, W, b, false, b, -1.0);
      g2.p(Ru + i * ab, a, b, a, 1.0).t(L, b, false, V, a, false, b, -1.0);
      g2.p(T, a, a, a, 1.0).t(R, b, false, V, a, false, b, -1.0);
    }
    if ((rc = g2.run(s))) return rc;
    const double* Bp = BDv(i - 1);
    Gq g3;
    g3.p(M1, b, b, b, 0.0).t(L, b, false, Bp, b, false, b, 1.0);
    if (a > 0) g3.p(M2, b, a, b, 0.0).t(R, b, false, Bp, b, false, b, 1.0);
    if ((rc = g3.run(s))) return rc;
    const double* BLp = BLov(i - 1);
    const double* BUp_ = BUpv(i - 1);
    const double* BRlp = BRl + (i - 1) * ab;
    const double* BRup = BRu + (i - 1) * ab;
    Gq g4;  // + M L^T - B_lower P^T
    g4.p(BDv(i), b, b, b, 1.0).t(M1, b, false, L, b, true, b, 1.0).t(BLp, b, false, P1, b, true, b, -1.0);
    if (a > 0) {
      g4.p(BRu + i * ab, a, b, a, 1.0).t(M1, b, false, R, b, true, b, 1.0).t(BLp, b, false, P2, b, true, b, -1.0);
      g4.p(BRl + i * ab, b, a, b, 1.0).t(M2, b, false, L, b, true, b, 1.0).t(BRlp, b, false, P1, b, true, b, -1.0);
      g4.p(BT, a, a, a, 1.0).t(M2, b, false, R, b, true, b, 1.0).t(BRlp, b, false, P2, b, true, b, -1.0);
    }
    if ((rc = g4.run(s))) return rc;
    Gq g5;  // - P B_upper
    g5.p(BDv(i), b, b, b, 1.0).t(P1, b, false, BUp_, b, false, b, -1.0);
    if (a > 0) {
      g5.p(BRu + i * ab, a, b, a, 1.0).t(P1, b, false, BRup, a, false, b, -1.0);
      g5.p(BRl + i * ab, b, a, b, 1.0).t(P2, b, false, BUp_, b, false, b, -1.0);
      g5.p(BT, a, a, a, 1.0).t(P2, b, false, BRup, a, false, b, -1.0);
    }
    if ((rc = g5.run(s))) return rc;
  }
  if (!invert_last_block) return ROCBTA_OK;
  double* Dl = Dv(n - 1);  // ddbtasc.py:447-475
  double* Bl = BDv(n - 1);
  rc = block_inverse(Dl, b, (int)b, ws, lay, s);
  if (rc) return rc;
  Gq h1;
  h1.p(Z, b, b, b, 0.0).t(Dl, b, false, Bl, b, false, b, 1.0);
  if (a > 0) h1.p(V, a, b, a, 0.0).t(Dl, b, false, Ru + (n - 1) * ab, a, false, b, 1.0);
  if ((rc = h1.run(s))) return rc;
  Gq h2;
  h2.p(Bl, b, b, b, 0.0).t(Z, b, false, Dl, b, true, b, 1.0);
  if (a > 0) h2.p(T, a, a, a, 1.0).t(Rl + (n - 1) * ab, b, false, V, a, false, b, -1.0);
  if ((rc = h2.run(s))) return rc;
  if (a == 0) return ROCBTA_OK;
  rc = block_inverse(T, a, (int)a, ws, lay, s);
  if (rc) return rc;
  const double* R = Rl + (n - 1) * ab;
  Gq h3;
  h3.p(P2, b, a, b, 0.0).t(R, b, false, Dl, b, false, b, 1.0);
  h3.p(M2, b, a, b, 0.0).t(R, b, false, Bl, b, false, b, 1.0);
  if ((rc = h3.run(s))) return rc;
  Gq h4;  // BT + R B' R^T - B_lower_arrow (R D)^T ...
  h4.p(BT, a, a, a, 1.0).t(M2, b, false, R, b, true, b, 1.0).t(BRl + (n - 1) * ab, b, false, P2, b, true, b, -1.0);
  if ((rc = h4.run(s))) return rc;
  Gq h5;  // ... - (R D) B_upper_arrow
  h5.p(BT, a, a, a, 1.0).t(P2, b, false, BRu + (n - 1) * ab, a, false, b, -1.0);
  if ((rc = h5.run(s))) return rc;
  Gq h6;  // T (..) T^T
  h6.p(Y1, a, a, a, 0.0).t(T, a, false, BT, a, false, a, 1.0);
  if ((rc = h6.run(s))) return rc;
  Gq h7;
  h7.p(BT, a, a, a, 0.0).t(Y1, a, false, T, a, true, a, 1.0);
  return h7.run(s);
}

static int dd_quadratic_sci(int64_t n, int64_t b, int64_t a, Bv Dv, Bv Lov, Bv Upv,
                            double* Rl, double* Ru, double* T, Bv BDv, Bv BLov, Bv BUpv,
                            double* BRl, double* BRu, double* BT, int invert_last_block,
                            void* workspace, size_t workspace_bytes, void* stream) {
  const DdLayout lay = dd_layout(b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  auto P = [&](size_t off) { return (double*)(ws + off); };
  double *B1 = P(lay.b1), *B2 = P(lay.b2), *C1 = P(lay.c1), *C2 = P(lay.c2), *E = P(lay.e);
  double *Gm = P(lay.g), *Dc = P(lay.dc), *Bdc = P(lay.bdc), *DF1 = P(lay.df1), *DF2 = P(lay.df2);
  double *DK1 = P(lay.dk1), *DK2 = P(lay.dk2), *S = P(lay.sm), *Mt = P(lay.mt), *Q1 = P(lay.q1);
  double *Q2 = P(lay.q2), *Q3 = P(lay.q3), *Y = P(lay.y), *V1 = P(lay.v1), *Zt = P(lay.zt);
  double *P3 = P(lay.p3), *D2 = P(lay.d2), *BB = P(lay.bbt), *Y2 = P(lay.y2);
  const int64_t bb = b * b, ab = a * b;
  int rc = ROCBTA_OK;
  if (invert_last_block && a > 0) {  // ddbtasci.py:377-430
    double* Ad = Dv(n - 1);
    double* Rul = Ru + (n - 1) * ab;
    double* Rll = Rl + (n - 1) * ab;
    double* Bd = BDv(n - 1);
    double* t13 = BRu + (n - 1) * ab;
    double* t31 = BRl + (n - 1) * ab;
    Gq m1;
    m1.p(B2, a, b, a, 0.0).t(Ad, b, false, Rul, a, false, b, 1.0);
    m1.p(Y, b, a, b, 0.0).t(Rll, b, false, Bd, b, false, b, 1.0);
    m1.p(V1, a, b, a, 0.0).t(Bd, b, false, Rll, b, true, b, 1.0);
    m1.p(Zt, a, b, a, 0.0).t(Ad, b, false, t13, a, false, b, 1.0);
    m1.p(P3, b, a, b, 0.0).t(T, a, false, t31, b, false, a, 1.0);
    m1.p(Y2, b, a, b, 0.0).t(Rll, b, false, Ad, b, false, b, 1.0);
    if ((rc = m1.run(s))) return rc;
    Gq m2;  // D2 = T Rl B, BB = B2 T
    m2.p(D2, b, a, b, 0.0).t(T, a, false, Y, b, false, a, 1.0);
    m2.p(BB, a, b, a, 0.0).t(B2, a, false, T, a, false, a, 1.0);
    if ((rc = m2.run(s))) return rc;
    Gq m3;
    m3.p(t13, a, b, a, 0.0).t(B2, a, false, BT, a, false, a, -1.0).t(V1, a, false, T, a, true, a, -1.0);
    m3.p(t31, b, a, b, 0.0).t(BT, a, false, B2, a, true, a, -1.0).t(P3, b, false, Ad, b, true, b, 1.0);
    if ((rc = m3.run(s))) return rc;
    Gq m4;
    m4.p(t13, a, b, a, 1.0).t(Zt, a, false, T, a, true, a, 1.0);
    m4.p(t31, b, a, b, 1.0).t(T, a, false, Y, b, false, a, -1.0);
    if ((rc = m4.run(s))) return rc;
    // B_D = B_D - B2 X^B_{t,n-1} + D2^T B2^T - Zt (B2 T)^T
    Gq m5;
    m5.p(Bd, b, b, b, 1.0).t(B2, a, false, t31, b, false, a, -1.0).t(D2, b, true, B2, a, true, a, 1.0);
    if ((rc = m5.run(s))) return rc;
    Gq m6;
    m6.p(Bd, b, b, b, 1.0).t(Zt, a, false, BB, a, true, a, -1.0);
    // A: X_{t,n-1} = -T (Rl Ad), X_{n-1,t} = -B2 T
    m6.p(Rll, b, a, b, 0.0).t(T, a, false, Y2, b, false, a, -1.0);
    m6.p(Rul, a, b, a, 0.0).t(B2, a, false, T, a, false, a, -1.0);
    if ((rc = m6.run(s))) return rc;
    Gq m7;
    m7.p(Ad, b, b, b, 1.0).t(B2, a, false, Rll, b, false, a, -1.0);
    if ((rc = m7.run(s))) return rc;
  }
  for (int64_t i = n - 2; i >= 0; --i) {  // ddbtasci.py:432-576
    double* Di = Dv(i);
    const double* D1 = Dv(i + 1);
    double* Li = Lov(i);
    double* Ui = Upv(i);
    double* Rli = Rl + i * ab;
    double* Rui = Ru + i * ab;
    const double* Rl1 = Rl + (i + 1) * ab;
    const double* Ru1 = Ru + (i + 1) * ab;
    double* Bi = BDv(i);
    const double* Bd1 = BDv(i + 1);
    const double* Bl1 = BRl + (i + 1) * ab;
    const double* Bu1 = BRu + (i + 1) * ab;
    double* b12 = BUpv(i);
    double* b13 = BRu + i * ab;
    double* b21 = BLov(i);
    double* b31 = BRl + i * ab;
    RB_CHECK(hipMemcpyAsync(Dc, Di, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    RB_CHECK(hipMemcpyAsync(Bdc, Bi, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    Gq la;
    la.p(B1, b, b, b, 0.0).t(Ui, b, false, D1, b, false, b, 1.0).t(Rui, a, false, Rl1, b, false, a, 1.0);
    la.p(C1, b, b, b, 0.0).t(D1, b, false, Li, b, false, b, 1.0).t(Ru1, a, false, Rli, b, false, a, 1.0);
    la.p(DF1, b, b, b, 0.0).t(Ui, b, false, Bd1, b, false, b, 1.0).t(Rui, a, false, Bl1, b, false, a, 1.0);
    la.p(DK1, b, b, b, 0.0).t(D1, b, false, b21, b, false, b, 1.0).t(Ru1, a, false, b31, b, false, a, 1.0);
    if (a > 0) {
      la.p(B2, a, b, a, 0.0).t(Ui, b, false, Ru1, a, false, b, 1.0).t(Rui, a, false, T, a, false, a, 1.0);
      la.p(C2, b, a, b, 0.0).t(Rl1, b, false, Li, b, false, b, 1.0).t(T, a, false, Rli, b, false, a, 1.0);
      la.p(DF2, a, b, a, 0.0).t(Ui, b, false, Bu1, a, false, b, 1.0).t(Rui, a, false, BT, a, false, a, 1.0);
      la.p(DK2, b, a, b, 0.0).t(Rl1, b, false, b21, b, false, b, 1.0).t(T, a, false, b31, b, false, a, 1.0);
    }
    if ((rc = la.run(s))) return rc;
    Gq lc;  // DF -= G, DK -= H
    lc.p(DF1, b, b, b, 1.0).t(b12, b, false, D1, b, true, b, -1.0).t(b13, a, false, Ru1, a, true, a, -1.0);
    lc.p(DK1, b, b, b, 1.0).t(Bd1, b, false, Ui, b, true, b, -1.0).t(Bu1, a, false, Rui, a, true, a, -1.0);
    if (a > 0) {
      lc.p(DF2, a, b, a, 1.0).t(b12, b, false, Rl1, b, true, b, -1.0).t(b13, a, false, T, a, true, a, -1.0);
      lc.p(DK2, b, a, b, 1.0).t(Bl1, b, false, Ui, b, true, b, -1.0).t(BT, a, false, Rui, a, true, a, -1.0);
    }
    if ((rc = lc.run(s))) return rc;
    Gq ld;
    ld.p(E, b, b, b, 0.0).t(B1, b, false, Li, b, false, b, 1.0).t(B2, a, false, Rli, b, false, a, 1.0);
    ld.p(S, b, b, b, 0.0).t(DF1, b, false, Ui, b, true, b, 1.0).t(DF2, a, false, Rui, a, true, a, 1.0);
    ld.p(Mt, b, b, b, 0.0).t(Ui, b, false, C1, b, false, b, 1.0).t(Rui, a, false, C2, b, false, a, 1.0);
    if ((rc = ld.run(s))) return rc;
    Gq le;
    le.p(S, b, b, b, 1.0).t(B1, b, false, b21, b, false, b, -1.0).t(B2, a, false, b31, b, false, a, -1.0);
    if ((rc = le.run(s))) return rc;
    Gq lf;
    lf.p(Q1, b, b, b, 0.0).t(Dc, b, false, S, b, false, b, 1.0);
    lf.p(Q2, b, b, b, 0.0).t(Dc, b, false, E, b, false, b, 1.0);
    lf.p(Q3, b, b, b, 0.0).t(Bdc, b, true, Mt, b, true, b, 1.0);
    if ((rc = lf.run(s))) return rc;
    Gq lg;
    lg.p(Bi, b, b, b, 1.0).t(Q1, b, false, Dc, b, true, b, 1.0).t(Q2, b, false, Bdc, b, false, b, 1.0);
    lg.p(b12, b, b, b, 0.0).t(Dc, b, false, DF1, b, false, b, -1.0).t(Bdc, b, false, C1, b, true, b, -1.0);
    lg.p(b21, b, b, b, 0.0).t(DK1, b, false, Dc, b, true, b, 1.0).t(C1, b, false, Bdc, b, false, b, -1.0);
    if (a > 0) {
      lg.p(b13, a, b, a, 0.0).t(Dc, b, false, DF2, a, false, b, -1.0).t(Bdc, b, false, C2, b, true, b, -1.0);
      lg.p(b31, b, a, b, 0.0).t(DK2, b, false, Dc, b, true, b, 1.0).t(C2, b, false, Bdc, b, false, b, -1.0);
    }
    if ((rc = lg.run(s))) return rc;
    Gq lh;
    lh.p(Bi, b, b, b, 1.0).t(Q3, b, false, Dc, b, true, b, 1.0);
    // A's selected inverse (ddbtasci.py:562-576)
    lh.p(Ui, b, b, b, 0.0).t(Dc, b, false, B1, b, false, b, -1.0);
    if (a > 0) lh.p(Rui, a, b, a, 0.0).t(Dc, b, false, B2, a, false, b, -1.0);
    if ((rc = lh.run(s))) return rc;
    Gq li;
    li.p(Li, b, b, b, 0.0).t(C1, b, false, Dc, b, false, b, -1.0);
    if (a > 0) li.p(Rli, b, a, b, 0.0).t(C2, b, false, Dc, b, false, b, -1.0);
    li.p(Gm, b, b, b, 0.0).t(Dc, b, false, E, b, false, b, 1.0);
    if ((rc = li.run(s))) return rc;
    Gq lj;
    lj.p(Di, b, b, b, 1.0).t(Gm, b, false, Dc, b, false, b, 1.0);
    if ((rc = lj.run(s))) return rc;
  }
  return ROCBTA_OK;
}

static int dd_quadratic_args(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                             double* Rl, double* Ru, double* T, double* BD, double* BLo,
                             double* BUp, double* BRl, double* BRu, double* BT, int direction) {
  if (bad_shape(n, b, a) || !D || !BD || (n > 1 && (!Lo || !Up || !BLo || !BUp)) ||
      (a > 0 && (!Rl || !Ru || !T || !BRl || !BRu || !BT)))
    return ROCBTA_ERR_ARG;
  if (direction != 0 && (direction != 1 || a > 0)) return ROCBTA_ERR_ARG;  // upward: BT only
  return ROCBTA_OK;
}

int rocbta_dddbtasc_quadratic(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                              double* Rl, double* Ru, double* T, double* BD, double* BLo,
                              double* BUp, double* BRl, double* BRu, double* BT,
                              int invert_last_block, int direction, void* workspace,
                              size_t workspace_bytes, void* stream) {
  int rc = dd_quadratic_args(n, b, a, D, Lo, Up, Rl, Ru, T, BD, BLo, BUp, BRl, BRu, BT, direction);
  if (rc) return rc;
  const int64_t bb = b * b;
  if (direction == 1)  // ddbtsc.py:314-356 as the downward sweep of J A J
    return dd_quadratic_sc(n, b, a, rev_view(D, n, bb), rev_view(Up, n - 1, bb),
                           rev_view(Lo, n - 1, bb), Rl, Ru, T, rev_view(BD, n, bb),
                           rev_view(BUp, n - 1, bb), rev_view(BLo, n - 1, bb), BRl, BRu, BT,
                           invert_last_block, workspace, workspace_bytes, stream);
  return dd_quadratic_sc(n, b, a, fwd_view(D, bb), fwd_view(Lo, bb), fwd_view(Up, bb), Rl, Ru, T,
                         fwd_view(BD, bb), fwd_view(BLo, bb), fwd_view(BUp, bb), BRl, BRu, BT,
                         invert_last_block, workspace, workspace_bytes, stream);
}

int rocbta_dddbtasci_quadratic(int64_t n, int64_t b, int64_t a, double* D, double* Lo,
                               double* Up, double* Rl, double* Ru, double* T, double* BD,
                               double* BLo, double* BUp, double* BRl, double* BRu, double* BT,
                               int invert_last_block, int direction, void* workspace,
                               size_t workspace_bytes, void* stream) {
  int rc = dd_quadratic_args(n, b, a, D, Lo, Up, Rl, Ru, T, BD, BLo, BUp, BRl, BRu, BT, direction);
  if (rc) return rc;
  const int64_t bb = b * b;
  if (direction == 1)  // ddbtsci.py:372-445 as the downward selected inversion of J A J
    return dd_quadratic_sci(n, b, a, rev_view(D, n, bb), rev_view(Up, n - 1, bb),
                            rev_view(Lo, n - 1, bb), Rl, Ru, T, rev_view(BD, n, bb),
                            rev_view(BUp, n - 1, bb), rev_view(BLo, n - 1, bb), BRl, BRu, BT,
                            invert_last_block, workspace, workspace_bytes, stream);
  return dd_quadratic_sci(n, b, a, fwd_view(D, bb), fwd_view(Lo, bb), fwd_view(Up, bb), Rl, Ru,
                          T, fwd_view(BD, bb), fwd_view(BLo, bb), fwd_view(BUp, bb), BRl, BRu, BT,
                          invert_last_block, workspace, workspace_bytes, stream);
}

// Permuted ("middle partition") forms of the non-SPD sweep: the middle
// blocks 1 .. n-2 are eliminated top-down; Lbuf[k] / Ubuf[k] carry the
// fill-ins A~_{0,k+1} / A~_{k+1,0} of row and column 0, and blocks 0, n-1
// and the tip are left as the reduced system.
static int dd_permuted_args(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                            double* Rl, double* Ru, double* T, double* Lbuf, double* Ubuf) {
  if (bad_shape(n, b, a) || n < 2 || !D || !Lo || !Up || !Lbuf || !Ubuf ||
      (a > 0 && (!Rl || !Ru || !T)))
    return ROCBTA_ERR_ARG;
  return ROCBTA_OK;
}

int rocbta_dddbtasc_permuted(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                             double* Rl, double* Ru, double* T, double* Lbuf, double* Ubuf,
                             void* workspace, size_t workspace_bytes, void* stream) {
  int rc = dd_permuted_args(n, b, a, D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf);
  if (rc) return rc;
  const DdLayout lay = dd_layout(b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double* W = (double*)(ws + lay.w);
  double* V = (double*)(ws + lay.v);
  double* Y = (double*)(ws + lay.z);
  const int64_t bb = b * b, ab = a * b;
  // ddbtasc.py:256-257: the fill-ins start as A_{0,1} and A_{1,0}
  RB_CHECK(hipMemcpyAsync(Lbuf, Up, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  RB_CHECK(hipMemcpyAsync(Ubuf, Lo, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  for (int64_t i = 1; i + 1 < n; ++i) {  // ddbtasc.py:259-332
    double* S = D + i * bb;
    const double* L = Lo + i * bb;
    const double* R = Rl + i * ab;
    const double* Lp = Lbuf + (i - 1) * bb;
    rc = block_inverse(S, b, (int)b, ws, lay, s);
    if (rc) return rc;
    // W = S^-1 A_{i,i+1}, Y = S^-1 A~_{i,0}, V = S^-1 A_{i,t}
    Gq g1;
    g1.p(W, b, b, b, 0.0).t(S, b, false, Up + i * bb, b, false, b, 1.0);
    g1.p(Y, b, b, b, 0.0).t(S, b, false, Ubuf + (i - 1) * bb, b, false, b, 1.0);
    if (a > 0) g1.p(V, a, b, a, 0.0).t(S, b, false, Ru + i * ab, a, false, b, 1.0);
    if ((rc = g1.run(s))) return rc;
    // every update reads only blocks this step does not write: one launch
    Gq g2;
    g2.p(D + (i + 1) * bb, b, b, b, 1.0).t(L, b, false, W, b, false, b, -1.0);
    g2.p(Lbuf + i * bb, b, b, b, 0.0).t(Lp, b, false, W, b, false, b, -1.0);
    g2.p(Ubuf + i * bb, b, b, b, 0.0).t(L, b, false, Y, b, false, b, -1.0);
    g2.p(D, b, b, b, 1.0).t(Lp, b, false, Y, b, false, b, -1.0);
    if (a > 0) {
      g2.p(Rl + (i + 1) * ab, b, a, b, 1.0).t(R, b, false, W, b, false, b, -1.0);
      g2.p(Ru + (i + 1) * ab, a, b, a, 1.0).t(L, b, false, V, a, false, b, -1.0);
      g2.p(T, a, a, a, 1.0).t(R, b, false, V, a, false, b, -1.0);
      g2.p(Rl, b, a, b, 1.0).t(R, b, false, Y, b, false, b, -1.0);
      g2.p(Ru, a, b, a, 1.0).t(Lp, b, false, V, a, false, b, -1.0);
    }
    if ((rc = g2.run(s))) return rc;
  }
  return ROCBTA_OK;
}

int rocbta_dddbtasci_permuted(int64_t n, int64_t b, int64_t a, double* D, double* Lo, double* Up,
                              double* Rl, double* Ru, double* T, double* Lbuf, double* Ubuf,
                              void* workspace, size_t workspace_bytes, void* stream) {
  int rc = dd_permuted_args(n, b, a, D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf);
  if (rc) return rc;
  const DdLayout lay = dd_layout(b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  auto P = [&](size_t off) { return (double*)(ws + off); };
  double *B1 = P(lay.b1), *B2 = P(lay.sm), *B3 = P(lay.b2), *C1 = P(lay.c1), *C2 = P(lay.mt);
  double *C3 = P(lay.c2), *E = P(lay.e), *Gm = P(lay.g), *Dc = P(lay.dc);
  const int64_t bb = b * b, ab = a * b;
  for (int64_t i = n - 2; i >= 1; --i) {  // ddbtasci.py:280-339
    double* Di = D + i * bb;
    const double* D1 = D + (i + 1) * bb;
    double* Li = Lo + i * bb;
    double* Ui = Up + i * bb;
    double* Rli = Rl + i * ab;
    double* Rui = Ru + i * ab;
    const double* Rl1 = Rl + (i + 1) * ab;
    const double* Ru1 = Ru + (i + 1) * ab;
    const double* Lb1 = Lbuf + i * bb;        // X_{0,i+1}
    const double* Ub1 = Ubuf + i * bb;        // X_{i+1,0}
    double* Lb0 = Lbuf + (i - 1) * bb;        // A~_{0,i} -> X_{0,i}
    double* Ub0 = Ubuf + (i - 1) * bb;        // A~_{i,0} -> X_{i,0}
    RB_CHECK(hipMemcpyAsync(Dc, Di, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    // B: row i of A~ times the finished columns; C: finished rows times column i
    const bool ar = a > 0;
    // B: row i of A~ times the finished columns; C: the finished rows times
    // column i.  Two terms per launch: the third (arrow) term follows.
    Gq g1;
    g1.p(B1, b, b, b, 0.0).t(Ui, b, false, D1, b, false, b, 1.0).t(Ub0, b, false, Lb1, b, false, b, 1.0);
    g1.p(B2, b, b, b, 0.0).t(Ui, b, false, Ub1, b, false, b, 1.0).t(Ub0, b, false, D, b, false, b, 1.0);
    g1.p(C1, b, b, b, 0.0).t(D1, b, false, Li, b, false, b, 1.0).t(Ub1, b, false, Lb0, b, false, b, 1.0);
    g1.p(C2, b, b, b, 0.0).t(Lb1, b, false, Li, b, false, b, 1.0).t(D, b, false, Lb0, b, false, b, 1.0);
    if (ar) {
      g1.p(B3, a, b, a, 0.0).t(Ui, b, false, Ru1, a, false, b, 1.0).t(Rui, a, false, T, a, false, a, 1.0);
      g1.p(C3, b, a, b, 0.0).t(Rl1, b, false, Li, b, false, b, 1.0).t(T, a, false, Rli, b, false, a, 1.0);
    }
    if ((rc = g1.run(s))) return rc;
    if (ar) {
      Gq h1;
      h1.p(B1, b, b, b, 1.0).t(Rui, a, false, Rl1, b, false, a, 1.0);
      h1.p(B2, b, b, b, 1.0).t(Rui, a, false, Rl, b, false, a, 1.0);
      h1.p(C1, b, b, b, 1.0).t(Ru1, a, false, Rli, b, false, a, 1.0);
      h1.p(C2, b, b, b, 1.0).t(Ru, a, false, Rli, b, false, a, 1.0);
      h1.p(B3, a, b, a, 1.0).t(Ub0, b, false, Ru, a, false, b, 1.0);
      h1.p(C3, b, a, b, 1.0).t(Rl, b, false, Lb0, b, false, b, 1.0);
      if ((rc = h1.run(s))) return rc;
    }
    // E = B1 A_{i+1,i} + B2 A~_{0,i} + B3 A~_{t,i} (old blocks); the new row i
    Gq g2;
    g2.p(E, b, b, b, 0.0).t(B1, b, false, Li, b, false, b, 1.0).t(B2, b, false, Lb0, b, false, b, 1.0);
    g2.p(Ui, b, b, b, 0.0).t(Dc, b, false, B1, b, false, b, -1.0);
    g2.p(Ub0, b, b, b, 0.0).t(Dc, b, false, B2, b, false, b, -1.0);
    if (ar) g2.p(Rui, a, b, a, 0.0).t(Dc, b, false, B3, a, false, b, -1.0);
    if ((rc = g2.run(s))) return rc;
    if (ar) {
      Gq h2;
      h2.p(E, b, b, b, 1.0).t(B3, a, false, Rli, b, false, a, 1.0);
      if ((rc = h2.run(s))) return rc;
    }
    // the new column i, and Gm = S^-1 E
    Gq g3;
    g3.p(Li, b, b, b, 0.0).t(C1, b, false, Dc, b, false, b, -1.0);
    g3.p(Lb0, b, b, b, 0.0).t(C2, b, false, Dc, b, false, b, -1.0);
    if (a > 0) g3.p(Rli, b, a, b, 0.0).t(C3, b, false, Dc, b, false, b, -1.0);
    g3.p(Gm, b, b, b, 0.0).t(Dc, b, false, E, b, false, b, 1.0);
    if ((rc = g3.run(s))) return rc;
    Gq g4;  // X_ii = S^-1 + S^-1 E S^-1
    g4.p(Di, b, b, b, 1.0).t(Gm, b, false, Dc, b, false, b, 1.0);
    if ((rc = g4.run(s))) return rc;
  }
  // ddbtasci.py:340-341: X_{1,0}, X_{0,1} back to the tridiagonal storage
  RB_CHECK(hipMemcpyAsync(Lo, Ubuf, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  RB_CHECK(hipMemcpyAsync(Up, Lbuf, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  return ROCBTA_OK;
}

// Sums of more than two products per output: the terms go out in rounds of
// two (a GEMM problem holds two), the first round with the output's beta
// and the later ones accumulating.  Outputs of one list must not be read by
// any term of the same list.
namespace {
struct Sums {
  struct Sum {
    double* C;
    int64_t ldc, M, N;
    double beta;
    std::vector<GemmTerm> terms;
  };
  std::vector<Sum> v;
  Sums& p(double* C, int64_t ldc, int64_t M, int64_t N, double beta) {
    v.push_back({C, ldc, M, N, beta, {}});
    return *this;
  }
  Sums& t(const double* A, int64_t lda, bool tA, const double* B, int64_t ldb, bool tB, int64_t k,
          double alpha) {
    if (k > 0) v.back().terms.push_back(term(A, lda, tA, B, ldb, tB, (int)k, alpha));
    return *this;
  }
  int run(hipStream_t s) {
    size_t rounds = 1;
    for (const Sum& x : v) rounds = std::max(rounds, (x.terms.size() + 1) / 2);
    for (size_t r = 0; r < rounds; ++r) {
      GemmList g;
      for (const Sum& x : v) {
        if (x.M <= 0 || x.N <= 0 || 2 * r >= x.terms.size()) continue;
        GemmProblem& pr = g.add(x.C, x.ldc, (int)x.M, (int)x.N, r == 0 ? x.beta : 1.0);
        for (size_t k = 2 * r; k < std::min(2 * r + 2, x.terms.size()); ++k) push_term(pr, x.terms[k]);
      }
      int rc = run_gemms(g, s);
      if (rc) return rc;
    }
    return ROCBTA_OK;
  }
};
// A block and its leading dimension.
struct Blk {
  double* p;
  int64_t ld;
};
}  // namespace

int rocbta_dddbtasc_quadratic_permuted(int64_t n, int64_t b, int64_t a, double* D, double* Lo,
                                       double* Up, double* Rl, double* Ru, double* T,
                                       double* Lbuf, double* Ubuf, double* BD, double* BLo,
                                       double* BUp, double* BRl, double* BRu, double* BT,
                                       double* BLbuf, double* BUbuf, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  int rc = dd_permuted_args(n, b, a, D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf);
  if (rc) return rc;
  if (!BD || !BLo || !BUp || !BLbuf || !BUbuf || (a > 0 && (!BRl || !BRu || !BT)))
    return ROCBTA_ERR_ARG;
  const DdLayout lay = dd_layout(b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  auto P = [&](size_t off) { return (double*)(ws + off); };
  double *W = P(lay.w), *V = P(lay.v), *Y = P(lay.z), *Z = P(lay.q1);
  const int64_t bb = b * b, ab = a * b;
  // ddbtasc.py:498-501: the fill-ins start as A's and B's (0,1) and (1,0)
  RB_CHECK(hipMemcpyAsync(Lbuf, Up, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  RB_CHECK(hipMemcpyAsync(Ubuf, Lo, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  RB_CHECK(hipMemcpyAsync(BLbuf, BUp, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  RB_CHECK(hipMemcpyAsync(BUbuf, BLo, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  for (int64_t i = 1; i + 1 < n; ++i) {  // ddbtasc.py:504-708
    double* S = D + i * bb;
    rc = block_inverse(S, b, (int)b, ws, lay, s);
    if (rc) return rc;
    // neighbours r = i+1, 0, t of block i: rows, F_r = A~_{r,i}, G_r =
    // B~_{r,i}, H_r = B~_{i,r}, P_r = F_r S, M_r = F_r S B~_ii S^T
    const int nr = a > 0 ? 3 : 2;
    const int64_t rows[3] = {b, b, a};
    const Blk F[3] = {{Lo + i * bb, b}, {Lbuf + (i - 1) * bb, b}, {Rl + i * ab, b}};
    const Blk G[3] = {{BLo + i * bb, b}, {BLbuf + (i - 1) * bb, b}, {BRl + i * ab, b}};
    const Blk H[3] = {{BUp + i * bb, b}, {BUbuf + (i - 1) * bb, b}, {BRu + i * ab, a}};
    const Blk Pm[3] = {{P(lay.p1), b}, {P(lay.m1), b}, {P(lay.p2), b}};
    const Blk Mm[3] = {{P(lay.q2), b}, {P(lay.q3), b}, {P(lay.m2), b}};
    // B~'s output block (r, c)
    const Blk Ob[3][3] = {{{BD + (i + 1) * bb, b}, {BUbuf + i * bb, b}, {BRu + (i + 1) * ab, a}},
                          {{BLbuf + i * bb, b}, {BD, b}, {BRu, a}},
                          {{BRl + (i + 1) * ab, b}, {BRl, b}, {BT, a}}};
    // W = S A_{i,i+1}, Y = S A~_{i,0}, V = S A_{i,t}; P_r; Z = S B~_ii
    Gq g1;
    g1.p(W, b, b, b, 0.0).t(S, b, false, Up + i * bb, b, false, b, 1.0);
    g1.p(Y, b, b, b, 0.0).t(S, b, false, Ubuf + (i - 1) * bb, b, false, b, 1.0);
    if (a > 0) g1.p(V, a, b, a, 0.0).t(S, b, false, Ru + i * ab, a, false, b, 1.0);
    for (int r = 0; r < nr; ++r)
      g1.p(Pm[r].p, Pm[r].ld, rows[r], b, 0.0).t(F[r].p, F[r].ld, false, S, b, false, b, 1.0);
    g1.p(Z, b, b, b, 0.0).t(S, b, false, BD + i * bb, b, false, b, 1.0);
    if ((rc = g1.run(s))) return rc;
    // A's updates (as rocbta_dddbtasc_permuted) and B~_ii <- Z S^T
    const double* L = Lo + i * bb;
    const double* R = Rl + i * ab;
    const double* Lp = Lbuf + (i - 1) * bb;
    Gq g2;
    g2.p(D + (i + 1) * bb, b, b, b, 1.0).t(L, b, false, W, b, false, b, -1.0);
    g2.p(Lbuf + i * bb, b, b, b, 0.0).t(Lp, b, false, W, b, false, b, -1.0);
    g2.p(Ubuf + i * bb, b, b, b, 0.0).t(L, b, false, Y, b, false, b, -1.0);
    g2.p(D, b, b, b, 1.0).t(Lp, b, false, Y, b, false, b, -1.0);
    if (a > 0) {
      g2.p(Rl + (i + 1) * ab, b, a, b, 1.0).t(R, b, false, W, b, false, b, -1.0);
      g2.p(Ru + (i + 1) * ab, a, b, a, 1.0).t(L, b, false, V, a, false, b, -1.0);
      g2.p(T, a, a, a, 1.0).t(R, b, false, V, a, false, b, -1.0);
      g2.p(Rl, b, a, b, 1.0).t(R, b, false, Y, b, false, b, -1.0);
      g2.p(Ru, a, b, a, 1.0).t(Lp, b, false, V, a, false, b, -1.0);
    }
    g2.p(BD + i * bb, b, b, b, 0.0).t(Z, b, false, S, b, true, b, 1.0);
    if ((rc = g2.run(s))) return rc;
    Gq g3;  // M_r = F_r B~_ii
    for (int r = 0; r < nr; ++r)
      g3.p(Mm[r].p, Mm[r].ld, rows[r], b, 0.0).t(F[r].p, F[r].ld, false, BD + i * bb, b, false, b, 1.0);
    if ((rc = g3.run(s))) return rc;
    // B~_rc += M_r F_c^T - G_r P_c^T - P_r H_c
    Sums g4;
    for (int r = 0; r < nr; ++r)
      for (int c = 0; c < nr; ++c)
        g4.p(Ob[r][c].p, Ob[r][c].ld, rows[r], rows[c], 1.0)
            .t(Mm[r].p, Mm[r].ld, false, F[c].p, F[c].ld, true, b, 1.0)
            .t(G[r].p, G[r].ld, false, Pm[c].p, Pm[c].ld, true, b, -1.0)
            .t(Pm[r].p, Pm[r].ld, false, H[c].p, H[c].ld, false, b, -1.0);
    if ((rc = g4.run(s))) return rc;
  }
  return ROCBTA_OK;
}

int rocbta_dddbtasci_quadratic_permuted(int64_t n, int64_t b, int64_t a, double* D, double* Lo,
                                        double* Up, double* Rl, double* Ru, double* T,
                                        double* Lbuf, double* Ubuf, double* BD, double* BLo,
                                        double* BUp, double* BRl, double* BRu, double* BT,
                                        double* BLbuf, double* BUbuf, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  int rc = dd_permuted_args(n, b, a, D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf);
  if (rc) return rc;
  if (!BD || !BLo || !BUp || !BLbuf || !BUbuf || (a > 0 && (!BRl || !BRu || !BT)))
    return ROCBTA_ERR_ARG;
  const DdLayout lay = dd_layout(b, a);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  auto P = [&](size_t off) { return (double*)(ws + off); };
  double *E = P(lay.e), *Sm = P(lay.sm), *Mt = P(lay.mt), *Q1 = P(lay.q1), *Q2 = P(lay.q2);
  double *Q3 = P(lay.q3), *Dc = P(lay.dc), *Bdc = P(lay.bdc), *Gm = P(lay.g);
  const int64_t bb = b * b, ab = a * b;
  const int nj = a > 0 ? 3 : 2;
  for (int64_t i = n - 2; i >= 1; --i) {  // ddbtasci.py:618-839
    // finished blocks j, k in (i+1, 0, t); sizes; A~'s row i (a_j) and
    // column i (c_j); B~'s (bR_j, bC_j); X_jk and XB_jk
    const int64_t sz[3] = {b, b, a};
    const Blk Aa[3] = {{Up + i * bb, b}, {Ubuf + (i - 1) * bb, b}, {Ru + i * ab, a}};
    const Blk Ac[3] = {{Lo + i * bb, b}, {Lbuf + (i - 1) * bb, b}, {Rl + i * ab, b}};
    const Blk bR[3] = {{BUp + i * bb, b}, {BUbuf + (i - 1) * bb, b}, {BRu + i * ab, a}};
    const Blk bC[3] = {{BLo + i * bb, b}, {BLbuf + (i - 1) * bb, b}, {BRl + i * ab, b}};
    const Blk X[3][3] = {{{D + (i + 1) * bb, b}, {Ubuf + i * bb, b}, {Ru + (i + 1) * ab, a}},
                         {{Lbuf + i * bb, b}, {D, b}, {Ru, a}},
                         {{Rl + (i + 1) * ab, b}, {Rl, b}, {T, a}}};
    const Blk XB[3][3] = {{{BD + (i + 1) * bb, b}, {BUbuf + i * bb, b}, {BRu + (i + 1) * ab, a}},
                          {{BLbuf + i * bb, b}, {BD, b}, {BRu, a}},
                          {{BRl + (i + 1) * ab, b}, {BRl, b}, {BT, a}}};
    // B_j (b x sz_j), C_j (sz_j x b), DF_j (b x sz_j), DK_j (sz_j x b)
    const Blk Bj[3] = {{P(lay.b1), b}, {P(lay.w), b}, {P(lay.b2), a}};
    const Blk Cj[3] = {{P(lay.c1), b}, {P(lay.z), b}, {P(lay.c2), b}};
    const Blk DF[3] = {{P(lay.df1), b}, {P(lay.p1), b}, {P(lay.df2), a}};
    const Blk DK[3] = {{P(lay.dk1), b}, {P(lay.m1), b}, {P(lay.dk2), b}};
    double* Di = D + i * bb;
    double* Bi = BD + i * bb;
    RB_CHECK(hipMemcpyAsync(Dc, Di, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    RB_CHECK(hipMemcpyAsync(Bdc, Bi, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
    // B_j = sum_k a_k X_kj; C_j = sum_k X_jk c_k; DF_j = sum_k a_k XB_kj -
    // sum_k bR_k X_jk^T; DK_j = sum_k X_jk bC_k - sum_k XB_jk a_k^T
    Sums pa;
    for (int j = 0; j < nj; ++j) {
      pa.p(Bj[j].p, Bj[j].ld, b, sz[j], 0.0);
      for (int k = 0; k < nj; ++k) pa.t(Aa[k].p, Aa[k].ld, false, X[k][j].p, X[k][j].ld, false, sz[k], 1.0);
      pa.p(Cj[j].p, Cj[j].ld, sz[j], b, 0.0);
      for (int k = 0; k < nj; ++k) pa.t(X[j][k].p, X[j][k].ld, false, Ac[k].p, Ac[k].ld, false, sz[k], 1.0);
      pa.p(DF[j].p, DF[j].ld, b, sz[j], 0.0);
      for (int k = 0; k < nj; ++k) {
        pa.t(Aa[k].p, Aa[k].ld, false, XB[k][j].p, XB[k][j].ld, false, sz[k], 1.0);
        pa.t(bR[k].p, bR[k].ld, false, X[j][k].p, X[j][k].ld, true, sz[k], -1.0);
      }
      pa.p(DK[j].p, DK[j].ld, sz[j], b, 0.0);
      for (int k = 0; k < nj; ++k) {
        pa.t(X[j][k].p, X[j][k].ld, false, bC[k].p, bC[k].ld, false, sz[k], 1.0);
        pa.t(XB[j][k].p, XB[j][k].ld, false, Aa[k].p, Aa[k].ld, true, sz[k], -1.0);
      }
    }
    if ((rc = pa.run(s))) return rc;
    // E = sum_j B_j c_j; Sm = sum_j DF_j a_j^T - sum_j B_j bC_j; Mt = sum_j a_j C_j
    Sums pb;
    pb.p(E, b, b, b, 0.0);
    for (int j = 0; j < nj; ++j) pb.t(Bj[j].p, Bj[j].ld, false, Ac[j].p, Ac[j].ld, false, sz[j], 1.0);
    pb.p(Sm, b, b, b, 0.0);
    for (int j = 0; j < nj; ++j) {
      pb.t(DF[j].p, DF[j].ld, false, Aa[j].p, Aa[j].ld, true, sz[j], 1.0);
      pb.t(Bj[j].p, Bj[j].ld, false, bC[j].p, bC[j].ld, false, sz[j], -1.0);
    }
    pb.p(Mt, b, b, b, 0.0);
    for (int j = 0; j < nj; ++j) pb.t(Aa[j].p, Aa[j].ld, false, Cj[j].p, Cj[j].ld, false, sz[j], 1.0);
    if ((rc = pb.run(s))) return rc;
    // Q1 = S Sm, Q2 = S E, Q3 = B~'^T Mt^T, Gm = S E; the new row and column
    // i of X B X^T (bR_j, bC_j) and of X (a_j, c_j)
    Gq pc;
    pc.p(Q1, b, b, b, 0.0).t(Dc, b, false, Sm, b, false, b, 1.0);
    pc.p(Q2, b, b, b, 0.0).t(Dc, b, false, E, b, false, b, 1.0);
    pc.p(Q3, b, b, b, 0.0).t(Bdc, b, true, Mt, b, true, b, 1.0);
    for (int j = 0; j < nj; ++j) {
      pc.p(bR[j].p, bR[j].ld, b, sz[j], 0.0)
          .t(Dc, b, false, DF[j].p, DF[j].ld, false, b, -1.0)
          .t(Bdc, b, false, Cj[j].p, Cj[j].ld, true, b, -1.0);
      pc.p(bC[j].p, bC[j].ld, sz[j], b, 0.0)
          .t(DK[j].p, DK[j].ld, false, Dc, b, true, b, 1.0)
          .t(Cj[j].p, Cj[j].ld, false, Bdc, b, false, b, -1.0);
      pc.p(Aa[j].p, Aa[j].ld, b, sz[j], 0.0).t(Dc, b, false, Bj[j].p, Bj[j].ld, false, b, -1.0);
      pc.p(Ac[j].p, Ac[j].ld, sz[j], b, 0.0).t(Cj[j].p, Cj[j].ld, false, Dc, b, false, b, -1.0);
    }
    if ((rc = pc.run(s))) return rc;
    Gq pd;  // X B X^T_ii = B~' + Q1 S^T + Q2 B~' (+ Q3 S^T below); X_ii = S + S E S
    pd.p(Bi, b, b, b, 1.0).t(Q1, b, false, Dc, b, true, b, 1.0).t(Q2, b, false, Bdc, b, false, b, 1.0);
    pd.p(Gm, b, b, b, 0.0).t(Dc, b, false, E, b, false, b, 1.0);
    if ((rc = pd.run(s))) return rc;
    Gq pe;
    pe.p(Bi, b, b, b, 1.0).t(Q3, b, false, Dc, b, true, b, 1.0);
    pe.p(Di, b, b, b, 1.0).t(Gm, b, false, Dc, b, false, b, 1.0);
    if ((rc = pe.run(s))) return rc;
  }
  // ddbtasci.py:841-845
  RB_CHECK(hipMemcpyAsync(Lo, Ubuf, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  RB_CHECK(hipMemcpyAsync(Up, Lbuf, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  RB_CHECK(hipMemcpyAsync(BLo, BUbuf, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  RB_CHECK(hipMemcpyAsync(BUp, BLbuf, bb * sizeof(double), hipMemcpyDeviceToDevice, s));
  return ROCBTA_OK;
}

// ------------------------------------------------------------- streaming
// Host <-> device streaming forms: the block arrays stay in host memory
// (pinned, for asynchronous copies), and two slots of device buffers per
// block kind carry the recurrence, with the copies on two streams of their
// own (host to device, device to host) overlapping the compute on the
// caller's stream.  For problems whose blocks do not fit in device memory.
// The calls return when the host arrays hold the results.
namespace {
struct CopyStreams {
  hipStream_t h2d = nullptr, d2h = nullptr;
  // ev[0..1]: h2d of slot k done; ev[2..3]: compute on slot k done;
  // ev[4..5]: d2h of slot k done
  hipEvent_t ev[6] = {};
  int init(hipStream_t s) {
    RB_CHECK(hipStreamCreateWithFlags(&h2d, hipStreamNonBlocking));
    RB_CHECK(hipStreamCreateWithFlags(&d2h, hipStreamNonBlocking));
    for (auto& e : ev) {
      RB_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      RB_CHECK(hipEventRecord(e, s));  // every wait below targets a recorded event
    }
    RB_CHECK(hipStreamWaitEvent(h2d, ev[0], 0));  // nothing starts before the caller's prior work
    RB_CHECK(hipStreamWaitEvent(d2h, ev[0], 0));
    return ROCBTA_OK;
  }
  // the caller's stream waits for both copy streams, then everything is done
  int finish(hipStream_t s) {
    hipEvent_t e1, e2;
    RB_CHECK(hipEventCreateWithFlags(&e1, hipEventDisableTiming));
    RB_CHECK(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
    RB_CHECK(hipEventRecord(e1, h2d));
    RB_CHECK(hipEventRecord(e2, d2h));
    RB_CHECK(hipStreamWaitEvent(s, e1, 0));
    RB_CHECK(hipStreamWaitEvent(s, e2, 0));
    RB_CHECK(hipStreamSynchronize(s));
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    return ROCBTA_OK;
  }
  // release (after the copy streams drained, also on an error return)
  ~CopyStreams() {
    if (h2d) (void)hipStreamSynchronize(h2d);
    if (d2h) (void)hipStreamSynchronize(d2h);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (h2d) (void)hipStreamDestroy(h2d);
    if (d2h) (void)hipStreamDestroy(d2h);
  }
};

int h2d_copy(double* dst, const double* src, int64_t count, hipStream_t s) {
  if (count <= 0) return ROCBTA_OK;
  RB_CHECK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyHostToDevice, s));
  return ROCBTA_OK;
}
int d2h_copy(double* dst, const double* src, int64_t count, hipStream_t s) {
  if (count <= 0) return ROCBTA_OK;
  RB_CHECK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, s));
  return ROCBTA_OK;
}
}  // namespace

int rocbta_dpobtaf_streaming(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                             double* T, int factorize_last_block, int* info, void* workspace,
                             size_t workspace_bytes, void* stream) {
  if (bad_shape(n, b, a) || !D || !info || (n > 1 && !L) || (a > 0 && (!R || !T)))
    return ROCBTA_ERR_ARG;
  const StLayout lay = st_layout(b, a, false);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double *Dd[2], *Ld[2], *Rd[2];
  for (int k = 0; k < 2; ++k) {
    Dd[k] = (double*)(ws + lay.d[k]);
    Ld[k] = (double*)(ws + lay.l[k]);
    Rd[k] = (double*)(ws + lay.r[k]);
  }
  double* Td = (double*)(ws + lay.t);
  double* tinv = (double*)(ws + lay.tinv);
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  set_int_kernel<<<1, 64, 0, s>>>(info, 0);
  RB_CHECK(hipGetLastError());
  CopyStreams cs;
  int rc = cs.init(s);
  if (rc) return rc;
  // pobtaf.py:343-362: the tip and block 0 in
  rc = h2d_copy(Td, T, a * a, cs.h2d);
  if (!rc) rc = h2d_copy(Dd[0], D, bb, cs.h2d);
  if (!rc) rc = h2d_copy(Rd[0], R, ab, cs.h2d);
  if (!rc && n > 1) rc = h2d_copy(Ld[0], L, bb, cs.h2d);
  if (rc) return rc;
  RB_CHECK(hipEventRecord(cs.ev[0], cs.h2d));
  RB_CHECK(hipStreamWaitEvent(s, cs.ev[0], 0));
  for (int64_t i = 0; i + 1 < n; ++i) {  // pobtaf.py:364-467
    const int k = (int)(i & 1), k1 = k ^ 1;
    // block i+1 into the other slot, once its previous contents are out
    RB_CHECK(hipStreamWaitEvent(cs.h2d, cs.ev[4 + k1], 0));
    rc = h2d_copy(Dd[k1], D + (i + 1) * bb, bb, cs.h2d);
    if (!rc) rc = h2d_copy(Rd[k1], R + (i + 1) * ab, ab, cs.h2d);
    if (!rc && i + 2 < n) rc = h2d_copy(Ld[k1], L + (i + 1) * bb, bb, cs.h2d);
    if (rc) return rc;
    RB_CHECK(hipEventRecord(cs.ev[k1], cs.h2d));
    // L_ii, L_{i+1,i}, L_{t,i}
    std::vector<Panel> panels = {{Ld[k], bi, b}};
    if (a > 0) panels.push_back({Rd[k], ai, b});
    rc = factor_block_column(Dd[k], b, bi, panels, i * b, info, tinv, s);
    if (rc) return rc;
    // Schur updates of block i+1 and the tip (pobtaf.py:443-467)
    RB_CHECK(hipStreamWaitEvent(s, cs.ev[k1], 0));
    const bool full_last = (i + 2 == n) && !factorize_last_block;
    GemmList upd;
    push_term(upd.add(Dd[k1], b, bi, bi, 1.0, full_last ? TRI_NONE : TRI_LOWER),
              term(Ld[k], b, false, Ld[k], b, true, bi, -1.0));
    if (a > 0) {
      push_term(upd.add(Rd[k1], b, ai, bi, 1.0), term(Rd[k], b, false, Ld[k], b, true, bi, -1.0));
      push_term(upd.add(Td, a, ai, ai, 1.0), term(Rd[k], b, false, Rd[k], b, true, bi, -1.0));
    }
    rc = run_gemms(upd, s);
    if (rc) return rc;
    RB_CHECK(hipEventRecord(cs.ev[2 + k], s));
    // factor blocks of step i out
    RB_CHECK(hipStreamWaitEvent(cs.d2h, cs.ev[2 + k], 0));
    rc = d2h_copy(D + i * bb, Dd[k], bb, cs.d2h);
    if (!rc) rc = d2h_copy(L + i * bb, Ld[k], bb, cs.d2h);
    if (!rc) rc = d2h_copy(R + i * ab, Rd[k], ab, cs.d2h);
    if (rc) return rc;
    RB_CHECK(hipEventRecord(cs.ev[4 + k], cs.d2h));
  }
  // last block and tip (pobtaf.py:469-522)
  const int kl = (int)((n - 1) & 1);
  if (factorize_last_block) {
    std::vector<Panel> panels;
    if (a > 0) panels.push_back({Rd[kl], ai, b});
    rc = factor_block_column(Dd[kl], b, bi, panels, (n - 1) * b, info, tinv, s);
    if (rc) return rc;
    if (a > 0) {
      GemmList upd;
      push_term(upd.add(Td, a, ai, ai, 1.0), term(Rd[kl], b, false, Rd[kl], b, true, bi, -1.0));
      rc = run_gemms(upd, s);
      if (!rc) rc = factor_block_column(Td, a, ai, {}, n * b, info, tinv, s);
      if (rc) return rc;
    }
  }
  RB_CHECK(hipEventRecord(cs.ev[2 + kl], s));
  RB_CHECK(hipStreamWaitEvent(cs.d2h, cs.ev[2 + kl], 0));
  rc = d2h_copy(D + (n - 1) * bb, Dd[kl], bb, cs.d2h);
  if (!rc) rc = d2h_copy(R + (n - 1) * ab, Rd[kl], ab, cs.d2h);
  if (!rc) rc = d2h_copy(T, Td, a * a, cs.d2h);
  if (rc) return rc;
  return cs.finish(s);
}

int rocbta_dpobtasi_streaming(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R,
                              double* T, int invert_last_block, void* workspace,
                              size_t workspace_bytes, void* stream) {
  if (bad_shape(n, b, a) || !D || (n > 1 && !L) || (a > 0 && (!R || !T))) return ROCBTA_ERR_ARG;
  const StLayout lay = st_layout(b, a, true);
  if (!workspace || workspace_bytes < lay.total) return ROCBTA_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)workspace;
  double *Dd[2], *Ld[2], *Rd[2], *XD[2], *XL[2], *XR[2];
  for (int k = 0; k < 2; ++k) {
    Dd[k] = (double*)(ws + lay.d[k]);
    Ld[k] = (double*)(ws + lay.l[k]);
    Rd[k] = (double*)(ws + lay.r[k]);
    XD[k] = (double*)(ws + lay.xd[k]);
    XL[k] = (double*)(ws + lay.xl[k]);
    XR[k] = (double*)(ws + lay.xr[k]);
  }
  double* Td = (double*)(ws + lay.t);
  double* Linv = (double*)(ws + lay.linv);
  double* W = (double*)(ws + lay.w);
  double* G = (double*)(ws + lay.g);
  double* Gt = (double*)(ws + lay.gt);
  double* Ti = (double*)(ws + lay.ti);
  double* Wt = (double*)(ws + lay.wt);
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  CopyStreams cs;
  int rc = cs.init(s);
  if (rc) return rc;
  const int kl = (int)((n - 1) & 1);
  // tip and last block in (pobtasi.py:318-333)
  rc = h2d_copy(Td, T, a * a, cs.h2d);
  if (!rc) rc = h2d_copy(invert_last_block ? Dd[kl] : XD[kl], D + (n - 1) * bb, bb, cs.h2d);
  if (!rc) rc = h2d_copy(invert_last_block ? Rd[kl] : XR[kl], R + (n - 1) * ab, ab, cs.h2d);
  if (rc) return rc;
  RB_CHECK(hipEventRecord(cs.ev[kl], cs.h2d));
  RB_CHECK(hipStreamWaitEvent(s, cs.ev[kl], 0));
  if (invert_last_block) {  // pobtasi.py:335-404
    if (a > 0) {
      rc = block_trtri(Td, a, 0, Ti, a, a * a, ai, 1, Wt, s);
      if (rc) return rc;
      GemmList xt;
      push_term(xt.add(Td, a, ai, ai, 0.0), term(Ti, a, true, Ti, a, false, ai, 1.0, TRI_UPPER,
                                                  TRI_LOWER));
      rc = run_gemms(xt, s);
      if (rc) return rc;
    }
    rc = block_trtri(Dd[kl], b, 0, Linv, b, bb, bi, 1, W, s, true);
    if (rc) return rc;
    GemmList g1;
    GemmProblem& ps = g1.add(XD[kl], b, bi, bi, 0.0, TRI_LOWER);
    ps.mirror = 1;
    push_term(ps, term(Linv, b, true, Linv, b, false, bi, 1.0, TRI_UPPER, TRI_LOWER));
    if (a > 0)
      push_term(g1.add(Gt, b, ai, bi, 0.0),
                term(Rd[kl], b, false, Linv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER));
    rc = run_gemms(g1, s);
    if (rc) return rc;
    if (a > 0) {
      GemmList g2;
      push_term(g2.add(XR[kl], b, ai, bi, 0.0), term(Td, a, false, Gt, b, false, ai, -1.0));
      rc = run_gemms(g2, s);
      if (rc) return rc;
      GemmList g3;
      GemmProblem& pd = g3.add(XD[kl], b, bi, bi, 1.0, TRI_LOWER);
      pd.mirror = 1;
      push_term(pd, term(XR[kl], b, true, Gt, b, false, ai, -1.0));
      rc = run_gemms(g3, s);
      if (rc) return rc;
    }
  }
  RB_CHECK(hipEventRecord(cs.ev[2 + kl], s));
  RB_CHECK(hipStreamWaitEvent(cs.d2h, cs.ev[2 + kl], 0));
  rc = d2h_copy(D + (n - 1) * bb, XD[kl], bb, cs.d2h);
  if (!rc) rc = d2h_copy(R + (n - 1) * ab, XR[kl], ab, cs.d2h);
  if (rc) return rc;
  RB_CHECK(hipEventRecord(cs.ev[4 + kl], cs.d2h));
  for (int64_t i = n - 2; i >= 0; --i) {  // pobtasi.py:406-494
    const int k = (int)(i & 1), k1 = k ^ 1;
    // factor blocks of step i in, once slot k's last compute (step i+2) is done
    RB_CHECK(hipStreamWaitEvent(cs.h2d, cs.ev[2 + k], 0));
    rc = h2d_copy(Dd[k], D + i * bb, bb, cs.h2d);
    if (!rc) rc = h2d_copy(Ld[k], L + i * bb, bb, cs.h2d);
    if (!rc) rc = h2d_copy(Rd[k], R + i * ab, ab, cs.h2d);
    if (rc) return rc;
    RB_CHECK(hipEventRecord(cs.ev[k], cs.h2d));
    RB_CHECK(hipStreamWaitEvent(s, cs.ev[k], 0));
    // the X blocks of slot k are rewritten: their copy-out (step i+2) first
    RB_CHECK(hipStreamWaitEvent(s, cs.ev[4 + k], 0));
    rc = block_trtri(Dd[k], b, 0, Linv, b, bb, bi, 1, W, s, true);
    if (rc) return rc;
    // G = L_{i+1,i} L_ii^{-1}, Gt = L_{t,i} L_ii^{-1}, X_ii <- L_ii^{-T} L_ii^{-1}
    GemmList g1;
    push_term(g1.add(G, b, bi, bi, 0.0),
              term(Ld[k], b, false, Linv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER));
    if (a > 0)
      push_term(g1.add(Gt, b, ai, bi, 0.0),
                term(Rd[k], b, false, Linv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER));
    GemmProblem& ps = g1.add(XD[k], b, bi, bi, 0.0, TRI_LOWER);
    ps.mirror = 1;
    push_term(ps, term(Linv, b, true, Linv, b, false, bi, 1.0, TRI_UPPER, TRI_LOWER));
    rc = run_gemms(g1, s);
    if (rc) return rc;
    // X_{i+1,i} = -X_{i+1,i+1} G - X_{t,i+1}^T Gt ; X_{t,i} = -X_{t,i+1} G - X_T Gt
    GemmList g2;
    GemmProblem& pl = g2.add(XL[k], b, bi, bi, 0.0);
    push_term(pl, term(XD[k1], b, false, G, b, false, bi, -1.0));
    if (a > 0) {
      push_term(pl, term(XR[k1], b, true, Gt, b, false, ai, -1.0));
      GemmProblem& pr = g2.add(XR[k], b, ai, bi, 0.0);
      push_term(pr, term(XR[k1], b, false, G, b, false, bi, -1.0));
      push_term(pr, term(Td, a, false, Gt, b, false, ai, -1.0));
    }
    rc = run_gemms(g2, s);
    if (rc) return rc;
    // X_ii += -X_{i+1,i}^T G - X_{t,i}^T Gt (symmetric: lower tiles, mirrored)
    GemmList g3;
    GemmProblem& pd = g3.add(XD[k], b, bi, bi, 1.0, TRI_LOWER);
    pd.mirror = 1;
    push_term(pd, term(XL[k], b, true, G, b, false, bi, -1.0));
    if (a > 0) push_term(pd, term(XR[k], b, true, Gt, b, false, ai, -1.0));
    rc = run_gemms(g3, s);
    if (rc) return rc;
    RB_CHECK(hipEventRecord(cs.ev[2 + k], s));
    RB_CHECK(hipStreamWaitEvent(cs.d2h, cs.ev[2 + k], 0));
    rc = d2h_copy(D + i * bb, XD[k], bb, cs.d2h);
    if (!rc) rc = d2h_copy(L + i * bb, XL[k], bb, cs.d2h);
    if (!rc) rc = d2h_copy(R + i * ab, XR[k], ab, cs.d2h);
    if (rc) return rc;
    RB_CHECK(hipEventRecord(cs.ev[4 + k], cs.d2h));
  }
  rc = d2h_copy(T, Td, a * a, cs.d2h);
  if (rc) return rc;
  return cs.finish(s);
}

}  // extern "C"

#include "batched.hpp"
