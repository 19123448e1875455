"""Benchmark: FP64 pobtaf + pobtasi on one MI355X (n=256, b=1024, a=32).

One step = restore the input blocks (device-to-device copy, untimed by the
events) + pobtaf + pobtasi on the seeded SPD BTA matrix, all resident in HBM.
`value` is the HIP-event time of pobtaf + pobtasi per step (mean over the
timed steps), in ms.  ms_per_step is the wall time per step of the bracketed
region (restore copy included).  pobtas (1 and 64 RHS) is timed separately.

The CPU baseline is the NumPy restatement of the reference algorithm
(oracle/serinv_np.py) timed on this host's cores on a sample of the same
workload (the first n_s block rows, b=1024, a=32), scaled to n=256 by the
algorithmic FLOP ratio.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N --b B --a A]
                       [--no-cpu-baseline] [--profile-steps K]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix peak (datasheet)
HBM_PEAK_GBS = 8000.0
METRIC = "pobtaf+pobtasi ms & %FP64-MFMA peak; mixed/batched speedup vs plain at stated err"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--b", type=int, default=1024)
    p.add_argument("--a", type=int, default=32)
    p.add_argument("--gamma", type=float, default=None,
                   help="diagonal-shift scale (default: calibrated value for kappa ~ 1e6)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-solve", action="store_true")
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--no-mixed", action="store_true")
    return p.parse_args()


def calibrated_gamma(n, b, a):
    path = os.path.join(ROOT, "serinv_amd", "utils", "calibration.json")
    try:
        with open(path) as f:
            table = json.load(f)
        return float(table[f"{n},{b},{a},{1e6:.0e}"]["gamma"]), "kappa~1e6 (calibrated)"
    except (OSError, KeyError, ValueError):
        return 1.0, "gamma=1 (strictly diagonally dominant)"


def measured_traffic(n, b, a):
    """HBM bytes per launch of pobtaf_chain_kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary for this shape (None if absent)."""
    path = os.path.join(ROOT, "profiles", f"traffic_pobtaf_chain_n{n}_b{b}_a{a}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_pobtaf_chain:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_baseline(b, a, n_full, budget_s):
    """Time the NumPy restatement on the first n_s block rows; scale by FLOPs."""
    import numpy as np
    from threadpoolctl import threadpool_info, threadpool_limits

    from oracle import serinv_np
    from serinv_amd.utils.matrices import bta_flops, bta_spd

    cores = len(os.sched_getaffinity(0))
    with threadpool_limits(limits=cores):
        info = threadpool_info()
        threads = max([i.get("num_threads", 1) for i in info] or [1])

        def run(ns):
            D, L, R, T, _ = bta_spd(ns, b, a, seed=63)
            t0 = time.perf_counter()
            serinv_np.pobtaf(D, L, R, T)
            t1 = time.perf_counter()
            serinv_np.pobtasi(D, L, R, T)
            t2 = time.perf_counter()
            return t1 - t0, t2 - t1

        run(2)  # warm-up
        ns = 4
        taf, tsi = run(ns)
        per = (taf + tsi) / max(sum(bta_flops(ns, b, a)), 1.0)
        ns2 = int(min(64, max(4, budget_s / max(per * sum(bta_flops(2, b, a)) / 2, 1e-9))))
        if ns2 > ns:
            ns = ns2
            taf, tsi = run(ns)
    f_s = sum(bta_flops(ns, b, a))
    f_full = sum(bta_flops(n_full, b, a))
    ms_full = (taf + tsi) * f_full / f_s * 1e3
    blas = ",".join(sorted({f"{i.get('internal_api')} {i.get('version')}" for i in info}))
    return {
        "value": round(ms_full, 1),
        "unit": "ms (pobtaf+pobtasi, scaled to the full workload)",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle/serinv_np.py pobtaf+pobtasi on n={ns} b={b} a={a} "
                  f"({taf + tsi:.2f} s measured: pobtaf {taf:.2f} s, pobtasi {tsi:.2f} s), "
                  f"scaled by FLOP ratio {f_full / f_s:.2f}; BLAS {blas}",
        "tflops": round(f_s / (taf + tsi) * 1e-12, 4),
    }


def bench_batched(dev, nmat=64, r=4, reps=3, tune=True):
    """64 independent SPD BTA systems (n ~ U{16..256}, b in {64, 128, 256},
    a ~ U{4..32}, seed 63 + k, 4 RHS): factor + solve + selected inversion,
    as one replayed batched plan (hipGraph) against a loop of the plain
    per-matrix calls.  Returns the timings and the largest difference
    between the two paths' outputs."""
    import numpy as np
    import torch

    import serinv_amd.algs as algs
    from serinv_amd.autotune import tune_batched
    from serinv_amd.batched import BatchedBTA
    from serinv_amd.utils.matrices import bta_flops, bta_solve_flops, bta_spd

    rng = np.random.default_rng(63)
    shapes = [(int(rng.integers(16, 257)), int(rng.choice([64, 128, 256])), int(rng.integers(4, 33)))
              for _ in range(nmat)]
    orig, work = [], []
    for k, (n, b, a) in enumerate(shapes):
        arrs = bta_spd(n, b, a, seed=63 + k, n_rhs=r)
        o = [torch.from_numpy(x).to(dev) for x in arrs]
        orig.append(o)
        work.append([t.clone() for t in o])
    flops = sum(sum(bta_flops(n, b, a)) + 2 * bta_solve_flops(n, b, a, r) for n, b, a in shapes)

    def restore():
        for w, o in zip(work, orig):
            for t, s in zip(w, o):
                t.copy_(s)
        torch.cuda.synchronize(dev)

    stream = torch.cuda.current_stream(dev)
    algs.config["check_info"] = False
    loop = []
    for _ in range(reps):
        restore()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for D, L, R, T, B in work:
            algs.pobtaf(D, L, R, T)
            algs.pobtas(D, L, R, T, B, trans="N")
            algs.pobtas(D, L, R, T, B, trans="C")
            algs.pobtasi(D, L, R, T)
        e1.record(stream)
        torch.cuda.synchronize(dev)
        loop.append(e0.elapsed_time(e1))
    loop_out = [[t.clone() for t in w] for w in work]
    tuned = None
    if tune:
        tuned = tune_batched([w[:4] for w in work], rhs=[w[4] for w in work], selinv=True,
                             use_cache=False)
    plan = BatchedBTA([w[:4] for w in work], rhs=[w[4] for w in work], selinv=True, graph=True)
    if tuned:
        plan.set_grid(tuned["factor_grid"], tuned["solve_grid"])
    restore()
    plan.run()  # capture + first replay
    torch.cuda.synchronize(dev)
    bat = []
    for _ in range(reps):
        restore()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        plan.run()
        e1.record(stream)
        torch.cuda.synchronize(dev)
        bat.append(e0.elapsed_time(e1))
    info, _, status = plan.results()
    diff = 0.0
    for w, lo in zip(work, loop_out):
        for t, s in zip(w, lo):
            if t.numel():
                diff = max(diff, float((t - s).abs().max() / s.abs().max().clamp_min(1e-300)))
    launches = plan.launch_count()
    plan.close()
    algs.config["check_info"] = True
    t_loop, t_bat = min(loop), min(bat)
    return {
        "workload": f"{nmat} SPD BTA systems, n~U[16,256], b in {{64,128,256}}, a~U[4,32], r={r}; "
                    "pobtaf + pobtas(N,C) + pobtasi, FP64",
        "batched_ms": round(t_bat, 3),
        "loop_ms": round(t_loop, 3),
        "speedup": round(t_loop / t_bat, 2),
        "matrices_per_s": round(nmat / (t_bat * 1e-3), 1),
        "tflops": round(flops / (t_bat * 1e-3) * 1e-12, 3),
        "launches_per_run": launches,
        "max_rel_diff_vs_loop": float(f"{diff:.3e}"),
        "info_ok": bool(np.all(info == 0) and status == 0),
        "autotune": tuned,
    }


def bench_mixed(dev, n, b, a, kappas=(1e4, 1e6, 1e8), r=16, reps=2):
    """Mixed FP32/FP64 against plain FP64 on the calibrated matrices of the
    given shape (cond ~ kappa): pobtaf + pobtasi time of both forms, max
    relative error of the mixed factor, selected inverse and log-determinant
    against the FP64 results, and the refined solve of r right-hand sides."""
    import numpy as np
    import torch

    import serinv_amd.algs as algs
    from serinv_amd.logdet import pobta_logdet
    from serinv_amd.mixed import pobtaf_mixed, pobtas_refine, pobtasi_mixed
    from serinv_amd.utils.conditioning import TABLE, ShiftedBTA

    with open(TABLE) as f:
        table = json.load(f)
    sh = ShiftedBTA(n, b, a, dev)
    stream = torch.cuda.current_stream(dev)
    rng = np.random.default_rng(64)
    Bh = torch.from_numpy(rng.uniform(-1.0, 1.0, (n * b + a, r))).to(dev)

    def relerr(x, ref):
        return float((x - ref).abs().max() / ref.abs().max())

    def timed(fa, fs, blocks):
        best = (float("inf"), 0.0, 0.0)
        for _ in range(reps):
            W = sh.blocks(gamma, [torch.empty_like(u) for u in sh.U]) if blocks is None else blocks
            torch.cuda.synchronize(dev)
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e[0].record(stream)
            fa(*W)
            e[1].record(stream)
            F = [x.clone() for x in W]
            e[2].record(stream)  # the clone is outside both timed ops
            e3 = torch.cuda.Event(enable_timing=True)
            e3.record(stream)
            fs(*W)
            e4 = torch.cuda.Event(enable_timing=True)
            e4.record(stream)
            torch.cuda.synchronize(dev)
            t = (e[0].elapsed_time(e[1]), e3.elapsed_time(e4))
            if sum(t) < best[0]:
                best = (sum(t), t[0], t[1], F, W)
        return best

    out = []
    algs.config["check_info"] = False
    for k in kappas:
        key = f"{n},{b},{a},{k:.0e}"
        if key not in table:
            continue
        gamma = table[key]["gamma"]
        A = sh.blocks(gamma)
        t64, af64, si64, F64, X64 = timed(algs.pobtaf, algs.pobtasi, None)
        t32, af32, si32, F32, X32 = timed(pobtaf_mixed, pobtasi_mixed, None)
        ld64 = float(pobta_logdet(F64[0], F64[3]))
        ld32 = float(pobta_logdet(F32[0], F32[3]))
        x64 = Bh.clone()
        algs.pobtas(*F64, x64, trans="N")
        algs.pobtas(*F64, x64, trans="C")
        xr = Bh.clone()
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        hist = pobtas_refine(A, F32, xr, max_iter=10, tol=1e-14)
        torch.cuda.synchronize(dev)
        t_ref = (time.perf_counter() - t0) * 1e3
        xm = Bh.clone()
        algs.pobtas(*F32, xm, trans="N")
        algs.pobtas(*F32, xm, trans="C")
        names = "DLRT"
        out.append({
            "kappa_target": k, "cond": round(table[key]["cond"], 1), "gamma": gamma,
            "fp64_ms": round(t64, 3), "mixed_ms": round(t32, 3),
            "fp64_pobtaf_ms": round(af64, 3), "mixed_pobtaf_ms": round(af32, 3),
            "fp64_pobtasi_ms": round(si64, 3), "mixed_pobtasi_ms": round(si32, 3),
            "speedup": round(t64 / t32, 3),
            "err_L": {nm: float(f"{relerr(x, y):.3e}") for nm, x, y in zip(names, F32, F64) if y.numel()},
            "err_X": {nm: float(f"{relerr(x, y):.3e}") for nm, x, y in zip(names, X32, X64) if y.numel()},
            "err_logdet": float(f"{abs(ld32 - ld64) / abs(ld64):.3e}"),
            "err_x_unrefined": float(f"{relerr(xm, x64):.3e}"),
            "err_x_refined": float(f"{relerr(xr, x64):.3e}"),
            "refine_residuals": [float(f"{h:.2e}") for h in hist],
            "refine_ms": round(t_ref, 2),
        })
        del A, F64, X64, F32, X32
        torch.cuda.empty_cache()
    algs.config["check_info"] = True
    return out


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the BTA factorization path here is single-GPU "
              "(see DESIGN.md for how it would shard); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import serinv_amd.algs as algs
    from serinv_amd.algs import config as algs_config
    from serinv_amd.utils.matrices import (bta_factor_bytes, bta_flops, bta_solve_flops,
                                           bta_spd)

    dev = torch.device("cuda:0")
    n, b, a = args.n, args.b, args.a
    gamma, cond_note = (args.gamma, f"gamma={args.gamma}") if args.gamma else calibrated_gamma(n, b, a)
    t0 = time.time()
    D, L, R, T, Bm = bta_spd(n, b, a, seed=63, gamma=gamma, n_rhs=64)
    print(f"[bench] generated n={n} b={b} a={a} gamma={gamma:g} in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    D0, L0, R0, T0 = (torch.from_numpy(x).to(dev) for x in (D, L, R, T))
    B0 = torch.from_numpy(Bm).to(dev)
    del D, L, R, T
    Dw, Lw, Rw, Tw = (torch.empty_like(x) for x in (D0, L0, R0, T0))
    stream = torch.cuda.current_stream(dev)
    algs_config["check_info"] = False

    def restore():
        for w, o in ((Dw, D0), (Lw, L0), (Rw, R0), (Tw, T0)):
            w.copy_(o)

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]

    def step(k=None):
        restore()
        if k is not None:
            ev[k][0].record(stream)
        algs.pobtaf(Dw, Lw, Rw, Tw)
        if k is not None:
            ev[k][1].record(stream)
        algs.pobtasi(Dw, Lw, Rw, Tw)
        if k is not None:
            ev[k][2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    t_start = time.perf_counter()
    for k in range(args.steps):
        step(k)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    t_af = [ev[k][0].elapsed_time(ev[k][1]) for k in range(args.steps)]
    t_si = [ev[k][1].elapsed_time(ev[k][2]) for k in range(args.steps)]
    t_tot = [x + y for x, y in zip(t_af, t_si)]
    ms = float(np.mean(t_tot))
    f_af, f_si = bta_flops(n, b, a)
    tflops = (f_af + f_si) / (ms * 1e-3) * 1e-12
    print(f"[bench] pobtaf {np.mean(t_af):.2f} ms  pobtasi {np.mean(t_si):.2f} ms  "
          f"total {ms:.2f} ms  {tflops:.2f} TFLOP/s", file=sys.stderr, flush=True)

    # Full-size property check on the last step's output: X_T from the
    # selected inversion against the tip of the solution A^{-1} E_t.
    extra = {}
    if not args.no_solve:
        algs_config["check_info"] = True
        restore()
        algs.pobtaf(Dw, Lw, Rw, Tw)
        torch.cuda.synchronize()
        sv = {}
        for r in (1, 64):
            Bw = B0[:, :r].contiguous().clone()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            algs.pobtas(Dw, Lw, Rw, Tw, Bw, trans="N")
            algs.pobtas(Dw, Lw, Rw, Tw, Bw, trans="C")
            e1.record(stream)
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1)
            sv[r] = (t, Bw)
            extra[f"pobtas_ms_r{r}"] = round(t, 3)
        byts = 2 * bta_factor_bytes(n, b, a)
        extra["pobtas_GBs_r1"] = round(byts / (sv[1][0] * 1e-3) * 1e-9, 1)
        # residual || A x - B || / ||B|| via block matvec on the original blocks
        x = sv[64][1]
        Ax = torch.zeros_like(x)
        xb = x[: n * b].view(n, b, -1)
        Ax[: n * b].view(n, b, -1).copy_(torch.bmm(D0, xb))
        if n > 1:
            Ax[b: n * b].view(n - 1, b, -1).add_(torch.bmm(L0, xb[:-1]))
            Ax[: (n - 1) * b].view(n - 1, b, -1).add_(torch.bmm(L0.transpose(1, 2).contiguous(), xb[1:]))
        if a:
            xt = x[n * b:]
            Rflat = R0.permute(1, 0, 2).contiguous().view(a, n * b)
            Ax[n * b:] += Rflat @ x[: n * b] + T0 @ xt
            Ax[: n * b].view(n, b, -1).add_(torch.bmm(R0.transpose(1, 2).contiguous(), xt.expand(n, a, -1).contiguous()))
        res = float((Ax - B0).norm() / B0.norm())
        extra["solve_rel_residual_r64"] = float(f"{res:.3e}")
        print(f"[bench] pobtas r=1 {extra['pobtas_ms_r1']} ms, r=64 {extra['pobtas_ms_r64']} ms, "
              f"residual {res:.2e}", file=sys.stderr, flush=True)

    # Roofline of the dominant kernel: pobtaf is one launch of the persistent
    # chain kernel (plus two flag fills of a few microseconds), so its event
    # time on this stream is that kernel's duration.
    tf_af = f_af / (float(np.mean(t_af)) * 1e-3) * 1e-12
    roofline = {
        "bound": "mfma",
        "achieved": round(tf_af, 3),
        "peak": FP64_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(tf_af / FP64_PEAK_TFLOPS, 4),
        "traffic": measured_traffic(n, b, a),
        "kernel": "pobtaf_chain_kernel (algorithmic pobtaf FLOPs / its HIP-event time)",
    }
    out = {
        "metric": METRIC,
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic seeded SPD BTA (seed 63, {cond_note})",
        "config": {"workload": f"FP64 pobtaf+pobtasi n={n} b={b} a={a}", "n": n, "b": b, "a": a,
                   "gamma": gamma},
        "pobtaf_ms": round(float(np.mean(t_af)), 3),
        "pobtasi_ms": round(float(np.mean(t_si)), 3),
        "tflops": round(tflops, 3),
        "frac_fp64_peak": round(tflops / FP64_PEAK_TFLOPS, 4),
        "pobtaf_tflops": round(tf_af, 3),
        "mixed_speedup": None,
        "batched_speedup": None,
        "roofline": roofline,
        **extra,
    }
    if not args.no_mixed:
        print("[bench] mixed FP32/FP64 vs FP64 ...", file=sys.stderr, flush=True)
        out["mixed"] = bench_mixed(dev, n, b, a)
        at6 = [m for m in out["mixed"] if m["kappa_target"] == 1e6]
        out["mixed_speedup"] = at6[0]["speedup"] if at6 else None
        for m in out["mixed"]:
            print(f"[bench] mixed kappa {m['kappa_target']:.0e}: {m['mixed_ms']} ms vs {m['fp64_ms']} ms, "
                  f"err X {max(m['err_X'].values()):.2e}, refined x {m['err_x_refined']:.2e}",
                  file=sys.stderr, flush=True)
    if not args.no_batched:
        print("[bench] batched plan vs per-matrix loop ...", file=sys.stderr, flush=True)
        out["batched"] = bench_batched(dev)
        out["batched_speedup"] = out["batched"]["speedup"]
        print(f"[bench] batched {out['batched']['batched_ms']} ms vs loop {out['batched']['loop_ms']} ms",
              file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(b, a, n, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
