"""Benchmark: FP64 pobtaf + pobtasi on one MI355X (n=256, b=1024, a=32).

One step = restore the input blocks (device-to-device copy, untimed by the
events) + pobtaf + pobtasi on the seeded SPD BTA matrix, all resident in HBM.
`value` is the HIP-event time of pobtaf + pobtasi per step (mean over the
timed steps), in ms.  ms_per_step is the wall time per step of the bracketed
region (restore copy included).  pobtas (1 and 64 RHS) is timed separately.

The CPU baseline is the NumPy restatement of the reference algorithm
(oracle/serinv_np.py) run once on the whole workload on this host's cores
(CPU-2, not scaled); it is also the parity oracle of the GPU outputs at this
size (`parity_vs_oracle`).  CPU-1 (n=32, b=64, a=16) is timed beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N --b B --a A]
                       [--no-cpu-baseline] [--no-solve] [--no-mixed] [--no-batched]
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
    p.add_argument("--no-solve", action="store_true")
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--no-mixed", action="store_true")
    p.add_argument("--no-permuted", action="store_true")
    p.add_argument("--no-streaming", action="store_true")
    p.add_argument("--no-ddbt", action="store_true")
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


def _cpu_info():
    import numpy as np
    from threadpoolctl import threadpool_info

    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info = threadpool_info()
    blas = ",".join(sorted({f"{i.get('internal_api')} {i.get('version')}" for i in info}))
    return model, blas, np.__version__


def cpu1_baseline(reps=5):
    """CPU-1: the oracle on n=32, b=64, a=16, r=1 (strictly diagonally
    dominant shifts), one warm-up, then the median of `reps` runs on fresh
    copies; per-op milliseconds.  Timed with one BLAS thread and with all
    cores; the faster setting is reported (64 x 64 blocks are too small for
    threaded BLAS calls to pay off)."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import serinv_np
    from serinv_amd.utils.matrices import bta_spd

    base = bta_spd(32, 64, 16, seed=63, n_rhs=1)
    best = None
    for threads in (1, len(os.sched_getaffinity(0))):
        times = []
        with threadpool_limits(limits=threads):
            for k in range(reps + 1):
                D, L, R, T, B = (x.copy() for x in base)
                t0 = time.perf_counter()
                serinv_np.pobtaf(D, L, R, T)
                t1 = time.perf_counter()
                serinv_np.pobtas(D, L, R, T, B, trans="N")
                serinv_np.pobtas(D, L, R, T, B, trans="C")
                t2 = time.perf_counter()
                serinv_np.pobtasi(D, L, R, T)
                t3 = time.perf_counter()
                if k:
                    times.append((t1 - t0, t3 - t2, t2 - t1))
        med = np.median(np.array(times), axis=0) * 1e3
        if best is None or med[0] + med[1] < best[1][0] + best[1][1]:
            best = (threads, med)
    threads, med = best
    return {"workload": "n=32 b=64 a=16 r=1", "pobtaf_ms": round(float(med[0]), 3),
            "pobtasi_ms": round(float(med[1]), 3), "pobtas_NC_ms": round(float(med[2]), 3),
            "runs": reps, "threads": threads}


def cpu2_oracle(host, gpu, n, b, a):
    """CPU-2: the oracle (oracle/serinv_np.py, the NumPy restatement of the
    reference routines) on the full workload, once, on all host cores.  It
    is the timed CPU baseline and the parity oracle of the GPU outputs at
    this size: every output of the GPU run in `gpu` is compared with it.
    Error metric per output array Y: max |Y - Y_ref| / max |Y_ref| over the
    stored blocks (relative |difference| for the log-determinant)."""
    import numpy as np
    from threadpoolctl import threadpool_info, threadpool_limits

    from oracle import serinv_np

    def rel(x, ref):
        return float(f"{np.max(np.abs(x - ref)) / np.max(np.abs(ref)):.3e}") if ref.size else 0.0

    cores = len(os.sched_getaffinity(0))
    D, L, R, T, Bm = host
    errs = {}
    with threadpool_limits(limits=cores):
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
        t0 = time.perf_counter()
        serinv_np.pobtaf(D, L, R, T)
        t_af = time.perf_counter() - t0
        errs["L"] = {nm: rel(g, o) for nm, g, o in zip("DLRT", gpu.pop("L"), (D, L, R, T))}
        ld = serinv_np.logdet_from_factor(D, T)
        errs["logdet"] = float(f"{abs(gpu['logdet'] - ld) / abs(ld):.3e}")
        t_sv = {}
        for r in (1, 64):
            x = np.ascontiguousarray(Bm[:, :r])
            t0 = time.perf_counter()
            serinv_np.pobtas(D, L, R, T, x, trans="N")
            serinv_np.pobtas(D, L, R, T, x, trans="C")
            t_sv[r] = time.perf_counter() - t0
            if f"x_r{r}" in gpu:
                errs[f"x_r{r}"] = rel(gpu.pop(f"x_r{r}"), x)
        t0 = time.perf_counter()
        serinv_np.pobtasi(D, L, R, T)
        t_si = time.perf_counter() - t0
        errs["X"] = {nm: rel(g, o) for nm, g, o in zip("DLRT", gpu.pop("X"), (D, L, R, T))}
    return t_af, t_si, t_sv, threads, errs


def bench_batched(dev, nmat=64, r=4, reps=3, tune=True, mixed=False):
    """64 independent SPD BTA systems (n ~ U{16..256}, b in {64, 128, 256},
    a ~ U{4..32}, seed 63 + k, 4 RHS): factor + solve + selected inversion,
    as one replayed batched plan (hipGraph) against a loop of the plain
    per-matrix calls.  Returns the timings and the largest difference
    between the two paths' outputs."""
    import numpy as np
    import torch

    import serinv_amd.algs as algs
    from serinv_amd.autotune import apply_tuning, tune_batched
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

    from serinv_amd.mixed import pobtaf_mixed, pobtasi_mixed

    factor = pobtaf_mixed if mixed else algs.pobtaf
    selinv = pobtasi_mixed if mixed else algs.pobtasi
    stream = torch.cuda.current_stream(dev)
    algs.config["check_info"] = False
    loop = []
    for _ in range(reps):
        restore()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for D, L, R, T, B in work:
            factor(D, L, R, T)
            algs.pobtas(D, L, R, T, B, trans="N")
            algs.pobtas(D, L, R, T, B, trans="C")
            selinv(D, L, R, T)
        e1.record(stream)
        torch.cuda.synchronize(dev)
        loop.append(e0.elapsed_time(e1))
    algs.check_status()
    loop_out = [[t.clone() for t in w] for w in work]
    tuned = None
    if tune:
        tuned = tune_batched([w[:4] for w in work], rhs=[w[4] for w in work], selinv=True,
                             use_cache=False, mixed=mixed)
    if tuned:
        plan = apply_tuning([w[:4] for w in work], tuned, rhs=[w[4] for w in work], selinv=True,
                            graph=True, mixed=mixed)
    else:
        plan = BatchedBTA([w[:4] for w in work], rhs=[w[4] for w in work], selinv=True, graph=True,
                          mixed=mixed)
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
                    "pobtaf + pobtas(N,C) + pobtasi, " + ("mixed FP32/FP64" if mixed else "FP64"),
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


def bench_modes(dev, n, b, a, a_wide=96):
    """The non-default pobtaf modes, one timed call each on n blocks:
    factorize_last_block=False (the distributed root's call, ppobtaf.py:
    110-118; on the chain kernel, with the last block and tip rebuilt) next to
    the default, and an arrow wider than one 64-tile (a_wide), which takes
    the launch-per-tile path.  The factorize_last_block=False result is
    checked against the default run's factor on the blocks both define."""
    import torch

    import serinv_amd.algs as algs
    from serinv_amd.utils.matrices import bta_spd

    stream = torch.cuda.current_stream(dev)

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return round(e0.elapsed_time(e1), 3)

    dev_of = lambda x: None if x is None else torch.from_numpy(x).to(dev)  # noqa: E731
    D0, L0, R0, T0, _ = (dev_of(x) for x in bta_spd(n, b, a, seed=63))
    res = {"n": n, "b": b, "a": a}
    D, L, R, T = (x.clone() for x in (D0, L0, R0, T0))
    res["pobtaf_ms"] = timed(lambda: algs.pobtaf(D, L, R, T))
    Do, Lo, Ro, To = (x.clone() for x in (D0, L0, R0, T0))
    res["pobtaf_open_last_ms"] = timed(lambda: algs.pobtaf(Do, Lo, Ro, To, factorize_last_block=False))
    diff = max(float((x[:-1] - y[:-1]).abs().max() / y[:-1].abs().max()) for x, y in ((Do, D), (Ro, R)))
    res["open_last_max_rel_diff_blocks_0_to_n-2"] = float(f"{diff:.3e}")
    del D, L, R, T, Do, Lo, Ro, To
    Dw, Lw, Rw, Tw, _ = (dev_of(x) for x in bta_spd(n, b, a_wide, seed=64))
    res["a_wide"] = a_wide
    res["pobtaf_a_wide_ms"] = timed(lambda: algs.pobtaf(Dw, Lw, Rw, Tw))
    return res


def bench_permuted(dev, n, b, a, reps=2):
    """One partition's work in the distributed form: the permuted ("middle
    partition") pobtaf and pobtasi on n blocks, completed through the 2-block
    reduced system, against the natural pobtaf + pobtasi of the same matrix
    (the first partition's form).  Both give the same selected inverse; the
    largest relative difference over D, L, R, T is reported."""
    import torch

    import serinv_amd.algs as algs
    from serinv_amd.utils import allocate_pobtax_permutation_buffers
    from serinv_amd.utils.matrices import bta_spd

    D0, L0, R0, T0, _ = (None if x is None else torch.from_numpy(x).to(dev)
                         for x in bta_spd(n, b, a, seed=63))
    stream = torch.cuda.current_stream(dev)

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    res = {"n": n, "b": b, "a": a}
    nat_af, nat_si, per_af, per_si = [], [], [], []
    for _ in range(reps):
        D, L, R, T = (x.clone() for x in (D0, L0, R0, T0))
        nat_af.append(timed(lambda: algs.pobtaf(D, L, R, T)))
        nat_si.append(timed(lambda: algs.pobtasi(D, L, R, T)))
        Dn, Ln, Rn, Tn = D, L, R, T
        D, L, R, T = (x.clone() for x in (D0, L0, R0, T0))
        buf = allocate_pobtax_permutation_buffers(D)
        per_af.append(timed(lambda: algs.pobtaf(D, L, R, T, buffer=buf)))
        # reduced system (block 0, block n-1, tip): natural factor + inverse
        rD = torch.stack([D[0], D[-1]]).contiguous()
        rL = buf[-1].T.contiguous().unsqueeze(0)
        rR = torch.stack([R[0], R[-1]]).contiguous()
        rT = T.clone()
        algs.pobtaf(rD, rL, rR, rT)
        algs.pobtasi(rD, rL, rR, rT)
        D[0], D[-1], buf[-1] = rD[0], rD[1], rL[0].T
        R[0], R[-1] = rR[0], rR[1]
        T.copy_(rT)
        per_si.append(timed(lambda: algs.pobtasi(D, L, R, T, buffer=buf)))
    torch.cuda.synchronize()

    def rel(x, y):
        return float((x - y).abs().max() / y.abs().max()) if y.numel() else 0.0

    res["natural_pobtaf_ms"] = round(min(nat_af), 3)
    res["natural_pobtasi_ms"] = round(min(nat_si), 3)
    res["permuted_pobtaf_ms"] = round(min(per_af), 3)
    res["permuted_pobtasi_ms"] = round(min(per_si), 3)
    res["max_rel_diff_vs_natural"] = float(f"{max(rel(D, Dn), rel(L, Ln), rel(R, Rn), rel(T, Tn)):.3e}")
    return res


def bench_streaming(dev, n, b, a, reps=2):
    """Host-streaming pobtaf + pobtasi (device_streaming=True, page-locked
    host blocks) against the in-memory calls on the same matrix; both give
    the same factor and selected inverse (largest relative difference
    reported).  Times include every host <-> device copy."""
    import time

    import torch

    import serinv_amd.algs as algs
    from serinv_amd.utils.matrices import bta_factor_bytes, bta_spd

    host = [torch.from_numpy(x).pin_memory() for x in bta_spd(n, b, a, seed=63)[:4]]
    res = {"n": n, "b": b, "a": a}
    t_af, t_si = [], []
    for _ in range(reps):
        D, L, R, T = (x.clone().pin_memory() for x in host)
        t0 = time.perf_counter()
        algs.pobtaf(D, L, R, T, device_streaming=True)
        t1 = time.perf_counter()
        algs.pobtasi(D, L, R, T, device_streaming=True)
        t2 = time.perf_counter()
        t_af.append(t1 - t0)
        t_si.append(t2 - t1)
    Dg, Lg, Rg, Tg = (x.to(dev) for x in host)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    algs.pobtaf(Dg, Lg, Rg, Tg)
    algs.pobtasi(Dg, Lg, Rg, Tg)
    torch.cuda.synchronize()
    t_mem = time.perf_counter() - t0

    def rel(x, y):
        y = y.cpu()
        return float((x - y).abs().max() / y.abs().max()) if y.numel() else 0.0

    res["streaming_pobtaf_ms"] = round(min(t_af) * 1e3, 3)
    res["streaming_pobtasi_ms"] = round(min(t_si) * 1e3, 3)
    res["in_memory_pobtaf_pobtasi_ms"] = round(t_mem * 1e3, 3)
    # each routine moves every block in and out once
    res["host_link_GBs_per_direction"] = round(
        bta_factor_bytes(n, b, a) / (min(t_af) + min(t_si)) * 2 * 1e-9, 1)
    res["max_rel_diff_vs_in_memory"] = float(
        f"{max(rel(D, Dg), rel(L, Lg), rel(R, Rg), rel(T, Tg)):.3e}")
    return res


def bench_ddbt(dev, n, b, a, reps=2):
    """Non-SPD family: ddbtasc + ddbtasci (and the A X A^T = B and permuted
    forms, one call each) on a diagonally dominant
    non-symmetric BTA matrix (random blocks in [0, 1), every diagonal entry
    raised by 3b + a + 1, generated on the device).  Property check: the
    diagonal blocks of A X are the identity, (A X)_ii = A_{i,i-1} X_{i-1,i}
    + A_ii X_ii + A_{i,i+1} X_{i+1,i} + A_{i,t} X_{t,i}."""
    import torch

    import serinv_amd.algs as algs

    g = torch.Generator(device=dev)
    g.manual_seed(63)
    rnd = lambda *shape: torch.rand(*shape, generator=g, dtype=torch.float64, device=dev)  # noqa: E731
    D0 = rnd(n, b, b) + (3 * b + a + 1) * torch.eye(b, dtype=torch.float64, device=dev)
    Lo0, Up0, Rl0, Ru0 = rnd(n - 1, b, b), rnd(n - 1, b, b), rnd(n, a, b), rnd(n, b, a)
    T0 = rnd(a, a) + (n * b + a + 1) * torch.eye(a, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    t_sc, t_sci = [], []
    for _ in range(reps):
        D, Lo, Up, Rl, Ru, T = (x.clone() for x in (D0, Lo0, Up0, Rl0, Ru0, T0))
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        e[0].record(stream)
        algs.ddbtasc(D, Lo, Up, Rl, Ru, T)
        e[1].record(stream)
        algs.ddbtasci(D, Lo, Up, Rl, Ru, T)
        e[2].record(stream)
        torch.cuda.synchronize()
        t_sc.append(e[0].elapsed_time(e[1]))
        t_sci.append(e[1].elapsed_time(e[2]))
    eye = torch.eye(b, dtype=torch.float64, device=dev)
    err = 0.0
    for i in range(n):
        ax = D0[i] @ D[i] + Ru0[i] @ Rl[i]
        if i > 0:
            ax = ax + Lo0[i - 1] @ Up[i - 1]
        if i + 1 < n:
            ax = ax + Up0[i] @ Lo[i]
        err = max(err, float((ax - eye).abs().max()))
    out = {"n": n, "b": b, "a": a, "ddbtasc_ms": round(min(t_sc), 3),
           "ddbtasci_ms": round(min(t_sci), 3), "max_abs_AX_ii_minus_I": float(f"{err:.3e}")}

    # the A X A^T = B form (B symmetric) and the permuted (middle-partition)
    # forms, one timed call each; the permuted pair is checked against the
    # natural selected inverse on the middle blocks' rows after the reduced
    # system (blocks 0, n-1, tip) is finished with the natural routines
    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return round(e0.elapsed_time(e1), 3)

    Bs = [x.clone() for x in (D0, Lo0, Up0, Rl0, Ru0, T0)]
    Bs[0] = 0.5 * (Bs[0] + Bs[0].transpose(1, 2))
    Bs[2] = Bs[1].transpose(1, 2).contiguous()
    Bs[4] = Bs[3].transpose(1, 2).contiguous()
    Bs[5] = 0.5 * (Bs[5] + Bs[5].T)
    keys = ("B_diagonal_blocks", "B_lower_diagonal_blocks", "B_upper_diagonal_blocks",
            "B_lower_arrow_blocks", "B_upper_arrow_blocks", "B_arrow_tip_block")
    A1 = [x.clone() for x in (D0, Lo0, Up0, Rl0, Ru0, T0)]
    rhs = dict(zip(keys, Bs))
    out["ddbtasc_quadratic_ms"] = timed(lambda: algs.ddbtasc(*A1, rhs=rhs, quadratic=True))
    out["ddbtasci_quadratic_ms"] = timed(lambda: algs.ddbtasci(*A1, rhs=rhs, quadratic=True))
    A2 = [x.clone() for x in (D0, Lo0, Up0, Rl0, Ru0, T0)]
    bufs = {"A_lower_buffer_blocks": torch.zeros_like(Lo0), "A_upper_buffer_blocks": torch.zeros_like(Lo0)}
    out["ddbtasc_permuted_ms"] = timed(lambda: algs.ddbtasc(*A2, buffers=bufs))
    # reduced system: the natural routines on blocks 0, n-1 and the tip
    red = [torch.stack([A2[0][0], A2[0][-1]]), bufs["A_upper_buffer_blocks"][-1:].clone(),
           bufs["A_lower_buffer_blocks"][-1:].clone(), torch.stack([A2[3][0], A2[3][-1]]),
           torch.stack([A2[4][0], A2[4][-1]]), A2[5].clone()]
    algs.ddbtasc(*red)
    algs.ddbtasci(*red)
    A2[0][0], A2[0][-1] = red[0][0], red[0][1]
    bufs["A_upper_buffer_blocks"][-1], bufs["A_lower_buffer_blocks"][-1] = red[1][0], red[2][0]
    A2[3][0], A2[3][-1], A2[4][0], A2[4][-1] = red[3][0], red[3][1], red[4][0], red[4][1]
    A2[5].copy_(red[5])
    out["ddbtasci_permuted_ms"] = timed(lambda: algs.ddbtasci(*A2, buffers=bufs))
    out["permuted_max_rel_diff_vs_natural"] = float(f"{float((A2[0] - D).abs().max() / D.abs().max()):.3e}")
    return out


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
    algs.check_status()
    algs.config["check_info"] = True
    return out


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the BTA factorization path here is single-GPU "
              "(the distributed form over several GPUs is not built); run with --gpus 1.",
              file=sys.stderr)
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
    host = (D, L, R, T, Bm)  # the oracle works on these in place, after the GPU runs
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
    algs.check_status()  # no pivot failure or scheduling time-out in the timed calls
    t_af = [ev[k][0].elapsed_time(ev[k][1]) for k in range(args.steps)]
    t_si = [ev[k][1].elapsed_time(ev[k][2]) for k in range(args.steps)]
    t_tot = [x + y for x, y in zip(t_af, t_si)]
    ms = float(np.mean(t_tot))
    f_af, f_si = bta_flops(n, b, a)
    tflops = (f_af + f_si) / (ms * 1e-3) * 1e-12
    print(f"[bench] pobtaf {np.mean(t_af):.2f} ms  pobtasi {np.mean(t_si):.2f} ms  "
          f"total {ms:.2f} ms  {tflops:.2f} TFLOP/s", file=sys.stderr, flush=True)

    # One more FP64 run of the three routines on the same inputs (checked
    # calls), whose outputs go to the host for the oracle comparison; the
    # solves are timed here (N then C, r = 1 and 64), and the residual
    # ||A x - B|| / ||B|| is formed on the GPU from the original blocks.
    extra = {}
    gpu_out = {}
    from serinv_amd.logdet import pobta_logdet
    from tests._blockprod import bta_matvec

    algs_config["check_info"] = True
    restore()
    algs.pobtaf(Dw, Lw, Rw, Tw)
    torch.cuda.synchronize()
    gpu_out["L"] = [x.cpu().numpy() for x in (Dw, Lw, Rw, Tw)]
    gpu_out["logdet"] = float(pobta_logdet(Dw, Tw))
    if not args.no_solve:
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
            res = float((bta_matvec((D0, L0, R0, T0), Bw) - B0[:, :r]).norm() / B0[:, :r].norm())
            extra[f"solve_rel_residual_r{r}"] = float(f"{res:.3e}")
            gpu_out[f"x_r{r}"] = Bw.cpu().numpy()
        byts = 2 * bta_factor_bytes(n, b, a)
        extra["pobtas_GBs_r1"] = round(byts / (sv[1][0] * 1e-3) * 1e-9, 1)
        print(f"[bench] pobtas r=1 {extra['pobtas_ms_r1']} ms, r=64 {extra['pobtas_ms_r64']} ms, "
              f"residuals {extra['solve_rel_residual_r1']:.2e} {extra['solve_rel_residual_r64']:.2e}",
              file=sys.stderr, flush=True)
        del sv
    algs.pobtasi(Dw, Lw, Rw, Tw)
    torch.cuda.synchronize()
    gpu_out["X"] = [x.cpu().numpy() for x in (Dw, Lw, Rw, Tw)]

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
    if not args.no_permuted:
        print("[bench] non-default pobtaf modes, n=64 ...", file=sys.stderr, flush=True)
        out["modes"] = bench_modes(dev, 64, b, a)
        print(f"[bench] modes {out['modes']}", file=sys.stderr, flush=True)
        print("[bench] permuted (middle-partition) form, n=64 ...", file=sys.stderr, flush=True)
        out["permuted"] = bench_permuted(dev, 64, b, a)
        print(f"[bench] permuted {out['permuted']}", file=sys.stderr, flush=True)
    if not args.no_streaming:
        print("[bench] host-streaming form, n=64 ...", file=sys.stderr, flush=True)
        out["streaming"] = bench_streaming(dev, 64, b, a)
        print(f"[bench] streaming {out['streaming']}", file=sys.stderr, flush=True)
    if not args.no_ddbt:
        print("[bench] non-SPD ddbtasc + ddbtasci, n=64 ...", file=sys.stderr, flush=True)
        out["ddbt"] = bench_ddbt(dev, 64, b, a)
        print(f"[bench] ddbt {out['ddbt']}", file=sys.stderr, flush=True)
    if not args.no_batched:
        print("[bench] batched plan vs per-matrix loop ...", file=sys.stderr, flush=True)
        out["batched"] = bench_batched(dev)
        out["batched_speedup"] = out["batched"]["speedup"]
        print(f"[bench] batched {out['batched']['batched_ms']} ms vs loop {out['batched']['loop_ms']} ms",
              file=sys.stderr, flush=True)
        out["batched_mixed"] = bench_batched(dev, mixed=True)
        print(f"[bench] batched mixed {out['batched_mixed']['batched_ms']} ms vs mixed loop "
              f"{out['batched_mixed']['loop_ms']} ms", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("[bench] CPU-1 and CPU-2: the oracle on the host ...", file=sys.stderr, flush=True)
        model, blas, npver = _cpu_info()
        cpu1 = cpu1_baseline()
        t_af, t_si, t_sv, threads, errs = cpu2_oracle(host, gpu_out, n, b, a)
        out["cpu_baseline"] = {
            "value": round((t_af + t_si) * 1e3, 1),
            "unit": "ms (pobtaf+pobtasi, full workload, one run)",
            "cores": threads,
            "kind": "port",
            "sample": f"oracle/serinv_np.py on the whole workload n={n} b={b} a={a} (CPU-2, one run, "
                      f"not scaled): pobtaf {t_af:.2f} s, pobtasi {t_si:.2f} s, pobtas N+C r=1 "
                      f"{t_sv.get(1, 0):.2f} s, r=64 {t_sv.get(64, 0):.2f} s; {model}; BLAS {blas}; "
                      f"numpy {npver}",
            "tflops": round((f_af + f_si) / (t_af + t_si) * 1e-12, 4),
            "cpu1": cpu1,
        }
        out["parity_vs_oracle"] = {
            "metric": "max|Y - Y_ref| / max|Y_ref| per output array (logdet: relative difference); "
                      "Y_ref from oracle/serinv_np.py on the same inputs (CPU-2)",
            **errs,
        }
        print(f"[bench] CPU-2 pobtaf {t_af:.2f} s pobtasi {t_si:.2f} s; GPU vs oracle {errs}",
              file=sys.stderr, flush=True)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
