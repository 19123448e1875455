// Batched plans: many independent SPD BTA systems of uneven n, b and a
// (included at the end of rocbta.hip; uses its op helpers).
//
// A plan is built once for a fixed set of device arrays and then replayed:
//   1. factorization of every matrix in ONE persistent launch
//      (pobtaf_chain_multi_kernel: one chain workgroup per matrix, a shared
//      pool of helper workgroups serving all matrices' task queues);
//   2. log-determinants (one workgroup per matrix);
//   3. optional solve x = A^{-1} B: one persistent multi-matrix sweep per
//      direction (pobtas_sweep_multi_kernel), reusing the diagonal-tile
//      inverses the factorization kernel left behind;
//   4. optional selected inversion: the per-matrix op sequence of pobtasi is
//      recorded, op k of every matrix merged into phase k, and each phase
//      runs as one fill / tile / grouped-GEMM launch over device-resident
//      descriptor lists.
// Every launch is a kernel (no memset / memcpy), so the whole sequence can
// be captured once into a hipGraph and replayed.
#pragma once

namespace {

__global__ void fill_list_kernel(const FillOp* ops) {
  const FillOp& op = ops[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < op.count;
       i += (int64_t)gridDim.x * blockDim.x)
    op.p[i] = op.v;
}

struct PhaseLaunch {
  int nfill = 0, ntile = 0, ngemm = 0;
  int64_t maxfill = 0;
  int tile_grid = 0, gemm_tiles = 0;
  size_t fill_off = 0, tile_off = 0, gemm_off = 0;  // offsets into the descriptor buffer
};

struct BatchedPlan {
  int dev = 0;
  int nmat = 0;
  int ops = 0;
  std::vector<int64_t> n, b, a, r;
  char* mem = nullptr;   // workspaces of all matrices
  char* desc = nullptr;  // descriptor lists
  // factorization
  PobtafChain* d_chains = nullptr;
  // The matrices run in waves of consecutive matrices, each wave one
  // persistent launch with at most max_chains chain workgroups, so that
  // helpers always have free CUs (factorization) or resident slots (sweeps)
  // whatever the batch size.  fwave[w] / swave[v]: first matrix of a wave.
  std::vector<int> fwave, swave;
  int* d_gclaim = nullptr;  // factor wave w: [w]; sweep wave v: [nfw + 2v + dir]
  int* d_cu_busy = nullptr;  // 4096 ints per factor wave
  int* d_info = nullptr;
  int* d_status = nullptr;
  FillOp* d_zero = nullptr;
  int nzero = 0;
  int64_t maxzero = 0;
  LogdetDesc* d_logdet = nullptr;
  double* d_logdet_out = nullptr;
  int chain_grid = 0;
  // solve
  SweepP* d_fwd = nullptr;
  SweepP* d_bwd = nullptr;
  int* d_chain_begin = nullptr;
  unsigned long long* d_empty = nullptr;
  int64_t nempty = 0;
  int sweep_rc = 1;
  int sweep_grid = 0;
  int gemm512 = 0;  // selected-inversion GEMM phases on 512-thread workgroups
  // selected inversion
  std::vector<PhaseLaunch> phases;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int record_pobtasi(int64_t n, int64_t b, int64_t a, double* D, double* L, double* R, double* T,
                   double* Linv, double* G, double* Gt, double* W, double* Tinv, double* Wt) {
  const int bi = (int)b, ai = (int)a;
  const int64_t bb = b * b, ab = a * b;
  hipStream_t s = nullptr;  // recording: nothing is launched
  // L_ii^{-1}; G_i = L_{i+1,i} L_ii^{-1}; Gt_i = L_{t,i} L_ii^{-1}; S_i -> D_i
  int rc = block_trtri(D, b, bb, Linv, b, bb, bi, (int)n, W, s);
  if (rc) return rc;
  GemmList pre;
  if (n > 1)
    push_term(pre.add(G, b, bi, bi, 0.0, TRI_NONE, (int)n - 1, bb),
              term(L, b, false, Linv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER, bb, bb));
  if (a > 0)
    push_term(pre.add(Gt, b, ai, bi, 0.0, TRI_NONE, (int)n, ab),
              term(R, b, false, Linv, b, false, bi, 1.0, TRI_NONE, TRI_LOWER, ab, bb));
  GemmProblem& psym = pre.add(D, b, bi, bi, 0.0, TRI_LOWER, (int)n, bb);
  psym.mirror = 1;
  push_term(psym, term(Linv, b, true, Linv, b, false, bi, 1.0, TRI_UPPER, TRI_LOWER, bb, bb));
  rc = run_gemms(pre, s);
  if (rc) return rc;
  if (a > 0) {  // last block and tip (pobtasi.py:114-139)
    double* Rl = R + (n - 1) * ab;
    double* Dl = D + (n - 1) * bb;
    rc = block_trtri(T, a, 0, Tinv, a, a * a, ai, 1, Wt, s);
    if (rc) return rc;
    GemmList xt;
    push_term(xt.add(T, a, ai, ai, 0.0),
              term(Tinv, a, true, Tinv, a, false, ai, 1.0, TRI_UPPER, TRI_LOWER));
    rc = run_gemms(xt, s);
    if (rc) return rc;
    GemmList xr;
    push_term(xr.add(Rl, b, ai, bi, 0.0), term(T, a, false, Gt + (n - 1) * ab, b, false, ai, -1.0));
    rc = run_gemms(xr, s);
    if (rc) return rc;
    GemmList xd;
    GemmProblem& pd = xd.add(Dl, b, bi, bi, 1.0, TRI_LOWER);
    pd.mirror = 1;
    push_term(pd, term(Rl, b, true, Gt + (n - 1) * ab, b, false, ai, -1.0));
    rc = run_gemms(xd, s);
    if (rc) return rc;
  }
  for (int64_t i = n - 2; i >= 0; --i) {  // pobtasi.py:141-172
    GemmList off;
    GemmProblem& pl = off.add(L + i * bb, b, bi, bi, 0.0);
    push_term(pl, term(D + (i + 1) * bb, b, false, G + i * bb, b, false, bi, -1.0));
    if (a > 0) {
      push_term(pl, term(R + (i + 1) * ab, b, true, Gt + i * ab, b, false, ai, -1.0));
      GemmProblem& pr = off.add(R + i * ab, b, ai, bi, 0.0);
      push_term(pr, term(R + (i + 1) * ab, b, false, G + i * bb, b, false, bi, -1.0));
      push_term(pr, term(T, a, false, Gt + i * ab, b, false, ai, -1.0));
    }
    rc = run_gemms(off, s);
    if (rc) return rc;
    GemmList dia;
    GemmProblem& pd = dia.add(D + i * bb, b, bi, bi, 1.0, TRI_LOWER);
    pd.mirror = 1;
    push_term(pd, term(L + i * bb, b, true, G + i * bb, b, false, bi, -1.0));
    if (a > 0) push_term(pd, term(R + i * ab, b, true, Gt + i * ab, b, false, ai, -1.0));
    rc = run_gemms(dia, s);
    if (rc) return rc;
  }
  return ROCBTA_OK;
}

template <int RC>
int launch_sweep_multi(const BatchedPlan& p, const SweepP* Ps, const int* chain_begin, int nm,
                       int* gclaim, hipStream_t s) {
  pobtas_sweep_multi_kernel<RC><<<p.sweep_grid, SW_THREADS, 0, s>>>(Ps, chain_begin, nm, gclaim);
  RB_CHECK(hipGetLastError());
  return ROCBTA_OK;
}

int batched_enqueue(const BatchedPlan& p, hipStream_t s) {
  const int64_t gx = std::min<int64_t>(1024, (p.maxzero + 255) / 256);
  fill_list_kernel<<<dim3((unsigned)std::max<int64_t>(gx, 1), (unsigned)p.nzero), 256, 0, s>>>(p.d_zero);
  RB_CHECK(hipGetLastError());
  const int nfw = (int)p.fwave.size() - 1;
  for (int w = 0; w < nfw; ++w) {
    const int m0 = p.fwave[w], nm = p.fwave[w + 1] - m0;
    pobtaf_chain_multi_kernel<<<std::max(nm + 1, p.chain_grid), CH_THREADS, 0, s>>>(
        p.d_chains + m0, nm, p.d_gclaim + w, p.d_cu_busy + (int64_t)w * 4096);
    RB_CHECK(hipGetLastError());
  }
  logdet_multi_kernel<<<p.nmat, 256, 0, s>>>(p.d_logdet);
  RB_CHECK(hipGetLastError());
  if (p.ops & ROCBTA_BATCHED_SOLVE) {
    const int nsw = (int)p.swave.size() - 1;
    for (int dir = 0; dir < 2; ++dir) {
      fill_u64_kernel<<<(int)std::min<int64_t>(4096, (p.nempty + 255) / 256), 256, 0, s>>>(
          p.d_empty, p.nempty, SW_EMPTY);
      RB_CHECK(hipGetLastError());
      for (int v = 0; v < nsw; ++v) {
        const int m0 = p.swave[v], nm = p.swave[v + 1] - m0;
        const SweepP* Ps = (dir == 0 ? p.d_fwd : p.d_bwd) + m0;
        const int* cb = p.d_chain_begin + m0 + v;  // per wave: nm + 1 relative entries
        int* gc = p.d_gclaim + nfw + 2 * v + dir;
        int rc = p.sweep_rc == 1 ? launch_sweep_multi<1>(p, Ps, cb, nm, gc, s)
                 : p.sweep_rc == 4 ? launch_sweep_multi<4>(p, Ps, cb, nm, gc, s)
                                   : launch_sweep_multi<8>(p, Ps, cb, nm, gc, s);
        if (rc) return rc;
      }
    }
  }
  if (p.ops & ROCBTA_BATCHED_SELINV) {
    for (const PhaseLaunch& ph : p.phases) {
      if (ph.nfill) {
        const int64_t fx = std::min<int64_t>(1024, (ph.maxfill + 255) / 256);
        fill_list_kernel<<<dim3((unsigned)fx, (unsigned)ph.nfill), 256, 0, s>>>(
            (const FillOp*)(p.desc + ph.fill_off));
        RB_CHECK(hipGetLastError());
      }
      if (ph.ntile) {
        tile_factor_list_kernel<<<ph.tile_grid, TT_THREADS, 0, s>>>(
            (const TileBatch*)(p.desc + ph.tile_off), ph.ntile);
        RB_CHECK(hipGetLastError());
      }
      if (ph.ngemm && ph.gemm_tiles > 0) {
        if (p.gemm512)
          gemm_list_kernel_h2<<<ph.gemm_tiles, 2 * GT_THREADS, 0, s>>>(
              (const GemmProblem*)(p.desc + ph.gemm_off), ph.ngemm);
        else
          gemm_list_kernel<<<ph.gemm_tiles, GT_THREADS, 0, s>>>(
              (const GemmProblem*)(p.desc + ph.gemm_off), ph.ngemm);
        RB_CHECK(hipGetLastError());
      }
    }
  }
  return ROCBTA_OK;
}

void batched_free(BatchedPlan* p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->mem) (void)hipFree(p->mem);
  if (p->desc) (void)hipFree(p->desc);
  delete p;
}

}  // namespace

extern "C" {

int rocbta_batched_create(int nmat, const int64_t* n, const int64_t* b, const int64_t* a,
                          double* const* D, double* const* L, double* const* R,
                          double* const* T, double* const* B, const int64_t* r, int ops,
                          rocbta_batched_plan* plan_out) {
  if (!plan_out || nmat < 1 || !n || !b || !a || !D || !L || !R || !T) return ROCBTA_ERR_ARG;
  *plan_out = nullptr;
  const bool solve = (ops & ROCBTA_BATCHED_SOLVE) != 0;
  if (solve && (!B || !r)) return ROCBTA_ERR_ARG;
  int64_t rmax = 1;
  for (int m = 0; m < nmat; ++m) {
    if (bad_shape(n[m], b[m], a[m]) || a[m] > CT || !D[m] || (n[m] > 1 && !L[m]) ||
        (a[m] > 0 && (!R[m] || !T[m])))
      return ROCBTA_ERR_ARG;
    if (solve && (r[m] < 1 || !B[m])) return ROCBTA_ERR_ARG;
    if (solve) rmax = std::max(rmax, r[m]);
  }
  // sweep chunk width: forced by the launch-shape bits, else by rmax
  const int rc_bits = ops & (ROCBTA_BATCHED_RC1 | ROCBTA_BATCHED_RC4 | ROCBTA_BATCHED_RC8);
  if (rc_bits & (rc_bits - 1)) return ROCBTA_ERR_ARG;  // at most one width
  const int rc_plan = rc_bits == ROCBTA_BATCHED_RC1   ? 1
                      : rc_bits == ROCBTA_BATCHED_RC4 ? 4
                      : rc_bits == ROCBTA_BATCHED_RC8 ? 8
                                                      : sweep_rc(rmax);
  // one matrix's column chunks must fit one sweep wave
  if (solve && cdiv(rmax, rc_plan) > SW_MAX_CHAINS) return ROCBTA_ERR_ARG;
  // chain_params and the recorded GEMM problems read the mixed flag
  MixedScope mixed((ops & ROCBTA_BATCHED_MIXED) != 0);
  BatchedPlan* p = new BatchedPlan();
  p->nmat = nmat;
  p->ops = ops;
  if (hipGetDevice(&p->dev) != hipSuccess) {
    delete p;
    return ROCBTA_ERR_HIP;
  }
  p->sweep_rc = rc_plan;
  p->gemm512 = (ops & ROCBTA_BATCHED_GEMM512) ? 1 : 0;
  const int RC = p->sweep_rc;
  // ---- layout of the workspace arena (offsets first, one allocation)
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align256(off + bytes);
    return o;
  };
  struct MatOff {
    size_t af, tipinv, si_linv, si_g, si_gt, si_w, si_tinv, si_wt, ybuf, part, apart;
    AfLayout lay;
    int nt, ta, NT, NG, nchunks;
  };
  std::vector<MatOff> mo(nmat);
  // waves: at most a quarter of the CUs run factorization chains at once,
  // and a sweep launch has at most SW_MAX_CHAINS chains
  const int cus = device_cus();
  const int fcap = std::max(1, cus / 4);
  for (int m = 0; m < nmat; m += fcap) p->fwave.push_back(m);
  p->fwave.push_back(nmat);
  if (solve) {
    int chains = 0;
    p->swave.push_back(0);
    for (int m = 0; m < nmat; ++m) {
      const int c = cdiv(r[m], RC);
      if (chains + c > SW_MAX_CHAINS) {
        p->swave.push_back(m);
        chains = 0;
      }
      chains += c;
    }
    p->swave.push_back(nmat);
  }
  const int nfw = (int)p->fwave.size() - 1, nsw = solve ? (int)p->swave.size() - 1 : 0;
  const size_t o_gclaim = take((size_t)(nfw + 2 * nsw) * sizeof(int));
  const size_t o_cu = take((size_t)nfw * 4096 * sizeof(int));
  const size_t o_info = take((size_t)nmat * sizeof(int));
  const size_t o_status = take((size_t)(1 + 2 * nmat) * sizeof(int));  // status, 2 counters per matrix
  const size_t o_ld = take((size_t)nmat * sizeof(double));
  for (int m = 0; m < nmat; ++m) {
    MatOff& q = mo[m];
    q.lay = af_layout(n[m], b[m], a[m]);
    q.af = take(q.lay.total);
    q.tipinv = take((size_t)CT * CT * sizeof(double));
    q.nt = cdiv(b[m], SW_T);
    q.ta = cdiv(a[m], SW_T);
    q.NT = (int)(n[m] * q.nt);
    q.NG = q.NT + q.ta;
    q.nchunks = solve ? cdiv(r[m], RC) : 0;
    if (ops & ROCBTA_BATCHED_SELINV) {
      const int64_t ntb = cdiv(b[m], TT), nta = cdiv(a[m], TT);
      q.si_linv = take((size_t)n[m] * b[m] * b[m] * sizeof(double));
      q.si_g = take((size_t)std::max<int64_t>(n[m] - 1, 0) * b[m] * b[m] * sizeof(double));
      q.si_gt = take((size_t)n[m] * a[m] * b[m] * sizeof(double));
      q.si_w = take((size_t)n[m] * ntb * TT * TT * sizeof(double));
      q.si_tinv = take((size_t)a[m] * a[m] * sizeof(double));
      q.si_wt = take((size_t)nta * TT * TT * sizeof(double));
    }
  }
  const size_t o_empty = off;  // sentinel-filled published blocks of all sweeps
  for (int m = 0; m < nmat && solve; ++m) {
    MatOff& q = mo[m];
    q.ybuf = take((size_t)q.nchunks * q.NG * SW_T * RC * sizeof(double));
    q.part = take((size_t)q.nchunks * q.NG * SW_T * RC * sizeof(double));
    q.apart = take((size_t)q.nchunks * n[m] * q.ta * SW_T * RC * sizeof(double));
  }
  const size_t o_empty_end = off;
  const size_t total = off;
  if (hipMalloc(&p->mem, total) != hipSuccess) {
    batched_free(p);
    return ROCBTA_ERR_HIP;
  }
  char* base = p->mem;
  p->d_gclaim = (int*)(base + o_gclaim);
  p->d_cu_busy = (int*)(base + o_cu);
  p->d_info = (int*)(base + o_info);
  p->d_status = (int*)(base + o_status);
  p->d_logdet_out = (double*)(base + o_ld);
  p->d_empty = (unsigned long long*)(base + o_empty);
  p->nempty = (int64_t)((o_empty_end - o_empty) / sizeof(double));

  // ---- descriptors (host), later copied to one device buffer
  std::vector<PobtafChain> chains(nmat);
  std::vector<LogdetDesc> lds(nmat);
  std::vector<SweepP> fwd, bwd;
  std::vector<int> chain_begin;  // per sweep wave: nm + 1 entries, relative to the wave
  std::vector<FillOp> zero;
  zero.push_back({(double*)(base + o_gclaim), (int64_t)((o_ld - o_gclaim) / sizeof(double)), 0.0});
  for (int m = 0; m < nmat; ++m) {
    MatOff& q = mo[m];
    char* ws = base + q.af;
    chains[m] = chain_params(n[m], b[m], a[m], D[m], L[m], R[m], T[m], p->d_info + m, ws, q.lay);
    chains[m].fused = 0;  // the many-matrix kernel runs the step-after-step chain
    chains[m].tip_inv = a[m] > 0 ? (double*)(base + q.tipinv) : nullptr;
    zero.push_back({(double*)(ws + q.lay.counter),
                    (int64_t)((q.lay.linv - q.lay.counter) / sizeof(double)), 0.0});
    lds[m] = {D[m], a[m] > 0 ? T[m] : nullptr, n[m], b[m], a[m], p->d_logdet_out + m};
    if (solve) {
      SweepP P{};
      P.D = D[m];
      P.L = L[m];
      P.R = R[m];
      P.T = T[m];
      P.B = B[m];
      P.ldb = r[m];
      P.n = (int)n[m];
      P.b = (int)b[m];
      P.a = (int)a[m];
      P.nt = q.nt;
      P.ta = q.ta;
      P.NT = q.NT;
      P.NG = q.NG;
      P.r = (int)r[m];
      P.rc = RC;
      P.nchunks = q.nchunks;
      P.tinv = chains[m].linv;
      P.tinv_tip = chains[m].tip_inv;
      P.ybuf = (double*)(base + q.ybuf);
      P.part = (double*)(base + q.part);
      P.apart = (double*)(base + q.apart);
      P.claim = nullptr;
      P.status = p->d_status;
      P.ntasks = 2 * q.NT * q.nchunks;
      P.spin_limit = 1ll << 22;
      P.trans = 0;
      P.counter = p->d_status + 1 + 2 * m;
      fwd.push_back(P);
      P.trans = 1;
      P.counter = p->d_status + 2 + 2 * m;
      bwd.push_back(P);
    }
  }
  for (int v = 0; solve && v + 1 < (int)p->swave.size(); ++v) {
    int c = 0;
    for (int m = p->swave[v]; m < p->swave[v + 1]; ++m) {
      chain_begin.push_back(c);
      c += mo[m].nchunks;
    }
    chain_begin.push_back(c);
  }
  p->nzero = (int)zero.size();
  for (const FillOp& z : zero) p->maxzero = std::max(p->maxzero, z.count);
  p->chain_grid = std::max(fcap + 1, 2 * cus);
  p->sweep_grid = std::max(SW_MAX_CHAINS + 1, 2 * cus);

  // ---- selected inversion: record every matrix's op sequence
  Recorder rec;
  if (ops & ROCBTA_BATCHED_SELINV) {
    g_rec = &rec;
    for (int m = 0; m < nmat; ++m) {
      rec.cur = 0;
      MatOff& q = mo[m];
      int rc = record_pobtasi(n[m], b[m], a[m], D[m], L[m], R[m], T[m],
                              (double*)(base + q.si_linv), (double*)(base + q.si_g),
                              (double*)(base + q.si_gt), (double*)(base + q.si_w),
                              (double*)(base + q.si_tinv), (double*)(base + q.si_wt));
      if (rc) {
        g_rec = nullptr;
        batched_free(p);
        return rc;
      }
    }
    g_rec = nullptr;
  }
  // ---- descriptor buffer
  size_t doff = 0;
  auto dtake = [&](size_t bytes) {
    size_t o = doff;
    doff = align256(doff + bytes);
    return o;
  };
  const size_t d_chains = dtake(nmat * sizeof(PobtafChain));
  const size_t d_lds = dtake(nmat * sizeof(LogdetDesc));
  const size_t d_zero = dtake(zero.size() * sizeof(FillOp));
  const size_t d_fwd = dtake(fwd.size() * sizeof(SweepP) + 1);
  const size_t d_bwd = dtake(bwd.size() * sizeof(SweepP) + 1);
  const size_t d_cb = dtake((chain_begin.size() + 1) * sizeof(int));
  for (Phase& ph : rec.phases) {
    PhaseLaunch pl;
    pl.nfill = (int)ph.fills.size();
    for (const FillOp& f : ph.fills) pl.maxfill = std::max(pl.maxfill, f.count);
    if (pl.nfill) pl.fill_off = dtake(ph.fills.size() * sizeof(FillOp));
    int grid = 0;
    for (TileBatch& tb : ph.tiles) {
      tb.grid_begin = grid;
      grid += tb.batch1 * tb.batch2;
    }
    pl.ntile = (int)ph.tiles.size();
    pl.tile_grid = grid;
    if (pl.ntile) pl.tile_off = dtake(ph.tiles.size() * sizeof(TileBatch));
    int64_t tiles = 0;
    for (GemmProblem& g : ph.gemms) {
      g.tile_begin = (int)tiles;
      tiles += (int64_t)g.tiles_m * g.tiles_n * g.batch1 * g.batch2;
    }
    pl.ngemm = (int)ph.gemms.size();
    pl.gemm_tiles = (int)tiles;
    if (pl.ngemm) pl.gemm_off = dtake(ph.gemms.size() * sizeof(GemmProblem));
    p->phases.push_back(pl);
  }
  if (hipMalloc(&p->desc, std::max<size_t>(doff, 256)) != hipSuccess) {
    batched_free(p);
    return ROCBTA_ERR_HIP;
  }
  std::vector<char> host(std::max<size_t>(doff, 256), 0);
  auto put = [&](size_t o, const void* src, size_t bytes) {
    if (bytes) memcpy(host.data() + o, src, bytes);
  };
  put(d_chains, chains.data(), nmat * sizeof(PobtafChain));
  put(d_lds, lds.data(), nmat * sizeof(LogdetDesc));
  put(d_zero, zero.data(), zero.size() * sizeof(FillOp));
  put(d_fwd, fwd.data(), fwd.size() * sizeof(SweepP));
  put(d_bwd, bwd.data(), bwd.size() * sizeof(SweepP));
  put(d_cb, chain_begin.data(), chain_begin.size() * sizeof(int));
  for (size_t k = 0; k < rec.phases.size(); ++k) {
    const Phase& ph = rec.phases[k];
    const PhaseLaunch& pl = p->phases[k];
    put(pl.fill_off, ph.fills.data(), ph.fills.size() * sizeof(FillOp));
    put(pl.tile_off, ph.tiles.data(), ph.tiles.size() * sizeof(TileBatch));
    put(pl.gemm_off, ph.gemms.data(), ph.gemms.size() * sizeof(GemmProblem));
  }
  if (hipMemcpy(p->desc, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
    batched_free(p);
    return ROCBTA_ERR_HIP;
  }
  p->d_chains = (PobtafChain*)(p->desc + d_chains);
  p->d_logdet = (LogdetDesc*)(p->desc + d_lds);
  p->d_zero = (FillOp*)(p->desc + d_zero);
  p->d_fwd = (SweepP*)(p->desc + d_fwd);
  p->d_bwd = (SweepP*)(p->desc + d_bwd);
  p->d_chain_begin = (int*)(p->desc + d_cb);
  p->n.assign(n, n + nmat);
  p->b.assign(b, b + nmat);
  p->a.assign(a, a + nmat);
  if (solve) p->r.assign(r, r + nmat);
  *plan_out = p;
  return ROCBTA_OK;
}

int rocbta_batched_set_grid(rocbta_batched_plan plan, int factor_grid, int solve_grid) {
  BatchedPlan* p = (BatchedPlan*)plan;
  if (!p) return ROCBTA_ERR_ARG;
  // every launch gets at least one workgroup more than its chains
  if (factor_grid > 0) p->chain_grid = factor_grid;
  if (solve_grid > 0) p->sweep_grid = std::max(SW_MAX_CHAINS + 1, solve_grid);
  if (p->exec) {  // the captured graph holds the old launch shapes
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  return ROCBTA_OK;
}

int rocbta_batched_run(rocbta_batched_plan plan, int use_graph, void* stream) {
  BatchedPlan* p = (BatchedPlan*)plan;
  if (!p) return ROCBTA_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (!use_graph) return batched_enqueue(*p, s);
  if (!p->exec) {
    RB_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = batched_enqueue(*p, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    RB_CHECK(e);
    p->graph = g;
    RB_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  }
  RB_CHECK(hipGraphLaunch(p->exec, s));
  return ROCBTA_OK;
}

int rocbta_batched_results(rocbta_batched_plan plan, int* info, double* logdet, int* status) {
  BatchedPlan* p = (BatchedPlan*)plan;
  if (!p) return ROCBTA_ERR_ARG;
  RB_CHECK(hipDeviceSynchronize());
  if (info) RB_CHECK(hipMemcpy(info, p->d_info, p->nmat * sizeof(int), hipMemcpyDeviceToHost));
  if (logdet)
    RB_CHECK(hipMemcpy(logdet, p->d_logdet_out, p->nmat * sizeof(double), hipMemcpyDeviceToHost));
  if (status) RB_CHECK(hipMemcpy(status, p->d_status, sizeof(int), hipMemcpyDeviceToHost));
  return ROCBTA_OK;
}

int rocbta_batched_launch_count(rocbta_batched_plan plan) {
  BatchedPlan* p = (BatchedPlan*)plan;
  if (!p) return -1;
  const int nfw = (int)p->fwave.size() - 1, nsw = (int)p->swave.size() - 1;
  int k = 2 + nfw + ((p->ops & ROCBTA_BATCHED_SOLVE) ? 2 * (1 + nsw) : 0);
  for (const PhaseLaunch& ph : p->phases)
    k += (ph.nfill > 0) + (ph.ntile > 0) + (ph.ngemm > 0 && ph.gemm_tiles > 0);
  return k;
}

int rocbta_batched_destroy(rocbta_batched_plan plan) {
  batched_free((BatchedPlan*)plan);
  return ROCBTA_OK;
}

}  // extern "C"
