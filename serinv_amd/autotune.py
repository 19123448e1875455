"""Autotuner for the batched plans' launch shapes.

The persistent launches of a batched plan (factorization and solve) take a
workgroup count; the best value depends on the mix of block sizes in the
batch: small b gives many short chains that want few helpers each, large b
long chains whose helper queues can keep more workgroups busy.  The plan
also has two tile shapes: the solve sweeps' column-chunk width (narrow
chunks give more, shorter chains; wide ones read each factor tile once for
more columns) and the workgroup size of the selected-inversion GEMM phases
(256 threads, two workgroups per CU, or 512 threads whose halves split a
tile's K range, which suits phases with few tiles).

tune_batched times the workgroup counts first (with the default shapes),
then the tile shapes with the best counts, on scratch copies of the batch
(the runs work in place), and returns the fastest; results are cached per
batch class (the multiset of (b, a-bucket) pairs, the RHS count and the
operations) in a JSON file.  apply_tuning builds the tuned plan.
"""

import json
import os
import time

import torch

from .batched import BatchedBTA

CACHE = os.environ.get("SERINV_AMD_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "serinv_amd", "autotune.json"))


def _cus(device):
    return torch.cuda.get_device_properties(device).multi_processor_count


def batch_class(systems, rhs=None):
    """Key of a batch: sorted counts of (b, a rounded up to 16), RHS width."""
    counts = {}
    for D, _, R, _ in systems:
        b, a = D.shape[1], R.shape[1]
        k = (b, -(-a // 16) * 16)
        counts[k] = counts.get(k, 0) + 1
    r = 0 if rhs is None else max(1 if x.dim() == 1 else x.shape[1] for x in rhs)
    return "r%d:" % r + ",".join(f"{b}x{a}*{c}" for (b, a), c in sorted(counts.items()))


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
    except OSError:
        pass


def shape_candidates(rhs=None, selinv=True):
    """Tile-shape candidates of a batch: (sweep_width, gemm_threads) pairs.

    sweep_width is None (the plan's default) without right-hand sides;
    gemm_threads is None without selected inversion."""
    widths = [None]
    if rhs is not None:
        r = max(1 if x.dim() == 1 else x.shape[1] for x in rhs)
        # at most 64 column chunks per system and no chunk wider than needed
        widths = [w for w in (1, 4, 8) if -(-r // w) <= 64 and (w == 1 or w < 2 * r)]
    threads = [256, 512] if selinv else [None]
    return [(w, t) for w in widths for t in threads]


def _time_plan(plan, systems, rhs, scratch, rscratch, reps, dev):
    best = float("inf")
    for _ in range(reps):
        for dst, src in zip(scratch, systems):
            for t, s in zip(dst, src):
                t.copy_(s)
        if rhs is not None:
            for t, s in zip(rscratch, rhs):
                t.copy_(s)
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        plan.run()
        torch.cuda.synchronize(dev)
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return round(best, 3)


def tune_batched(systems, rhs=None, selinv=True, candidates=None, reps=2, use_cache=True,
                 mixed=False, shapes=None):
    """Return {"factor_grid", "solve_grid", "sweep_width", "gemm_threads", "ms",
    "trials": [...]} for the fastest launch shape of this batch."""
    dev = systems[0][0].device
    key = batch_class(systems, rhs) + (":si" if selinv else "") + (":mx" if mixed else "")
    table = _load() if use_cache else {}
    if key in table and "sweep_width" in table[key]:
        return table[key]
    cus = _cus(dev)
    if candidates is None:
        candidates = [(f * cus, s * cus) for f in (1, 2, 3) for s in (1, 2)]
    if shapes is None:
        shapes = shape_candidates(rhs, selinv)
    scratch = [[t.clone() for t in sysm] for sysm in systems]
    rscratch = None if rhs is None else [x.clone() for x in rhs]
    trials = []
    # 1. workgroup counts with the default shapes
    plan = BatchedBTA(scratch, rhs=rscratch, selinv=selinv, graph=False, mixed=mixed)
    try:
        for fg, sg in candidates:
            plan.set_grid(fg, sg)
            trials.append({"factor_grid": fg, "solve_grid": sg, "sweep_width": None,
                           "gemm_threads": None,
                           "ms": _time_plan(plan, systems, rhs, scratch, rscratch, reps, dev)})
    finally:
        plan.close()
    top = min(trials, key=lambda t: t["ms"])
    # 2. tile shapes with the best counts (a shape is fixed when the plan is built)
    for w, th in shapes:
        if w is None and th in (None, 256):
            continue  # the default shape, timed above
        plan = BatchedBTA(scratch, rhs=rscratch, selinv=selinv, graph=False, mixed=mixed,
                          sweep_width=w, gemm_threads=th)
        try:
            plan.set_grid(top["factor_grid"], top["solve_grid"])
            trials.append({"factor_grid": top["factor_grid"], "solve_grid": top["solve_grid"],
                           "sweep_width": w, "gemm_threads": th,
                           "ms": _time_plan(plan, systems, rhs, scratch, rscratch, reps, dev)})
        finally:
            plan.close()
    top = min(trials, key=lambda t: t["ms"])
    out = dict(top, trials=trials)
    if use_cache:
        table[key] = out
        _store(table)
    return out


def apply_tuning(systems, tuning, rhs=None, selinv=True, graph=True, mixed=False):
    """A BatchedBTA plan with the launch shape tune_batched returned."""
    plan = BatchedBTA(systems, rhs=rhs, selinv=selinv, graph=graph, mixed=mixed,
                      sweep_width=tuning.get("sweep_width"), gemm_threads=tuning.get("gemm_threads"))
    plan.set_grid(tuning.get("factor_grid", 0), tuning.get("solve_grid", 0))
    return plan
