"""Autotuner for the batched plans' launch shapes.

The persistent launches of a batched plan (factorization and solve) take a
workgroup count; the best value depends on the mix of block sizes in the
batch: small b gives many short chains that want few helpers each, large b
long chains whose helper queues can keep more workgroups busy.  tune_batched
times every candidate on scratch copies of the batch (the runs work in
place) and returns the fastest; results are cached per batch class
(the multiset of (b, a-bucket) pairs and the RHS count) in a JSON file.
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


def tune_batched(systems, rhs=None, selinv=True, candidates=None, reps=2, use_cache=True,
                 mixed=False):
    """Return {"factor_grid": g1, "solve_grid": g2, "ms": t, "trials": [...]}."""
    dev = systems[0][0].device
    key = batch_class(systems, rhs) + (":si" if selinv else "") + (":mx" if mixed else "")
    table = _load() if use_cache else {}
    if key in table:
        return table[key]
    cus = _cus(dev)
    if candidates is None:
        candidates = [(f * cus, s * cus) for f in (1, 2, 3) for s in (1, 2)]
    scratch = [[t.clone() for t in sysm] for sysm in systems]
    rscratch = None if rhs is None else [x.clone() for x in rhs]
    plan = BatchedBTA(scratch, rhs=rscratch, selinv=selinv, graph=False, mixed=mixed)
    trials = []
    try:
        for fg, sg in candidates:
            plan.set_grid(fg, sg)
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
            trials.append({"factor_grid": fg, "solve_grid": sg, "ms": round(best, 3)})
    finally:
        plan.close()
    top = min(trials, key=lambda t: t["ms"])
    out = dict(top, trials=trials)
    if use_cache:
        table[key] = out
        _store(table)
    return out
