"""Condition-number calibration of the seeded SPD BTA generator.

bta_spd(n, b, a, gamma) builds A(gamma) = U + gamma * S, where U holds the
seeded uniform[0,1) blocks and S is diagonal (3b + a on the block rows,
n*b + a on the tip rows).  Smaller gamma moves A towards singularity.  For a
target cond_2(A) this module finds gamma on the GPU:

* the PD boundary gamma0 by bisection on the success of the factorization
  (numpy.linalg.LinAlgError from pobtaf = not positive definite);
* lambda_max by power iteration with the FP64 BTA product (rocbta_dresidual
  with B = 0 gives -A x), lambda_min by inverse iteration with the factor
  (pobtas forward + backward), cond = lambda_max / lambda_min;
* gamma for the target by bisection on log(gamma - gamma0).

calibrate() stores the result in calibration.json next to this file, keyed
"n,b,a,kappa"; the benchmark reads it so that every run uses the same
matrices.  All of this runs the library's own kernels; it is a tool, not part
of the solve path.
"""

import json
import math
import os

import numpy as np
import torch

from .matrices import bta_spd

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "calibration.json")


class ShiftedBTA:
    """U (the gamma = 0 blocks) on the device; blocks(gamma) = U + gamma * S."""

    def __init__(self, n, b, a, device, seed=63):
        D, L, R, T, _ = bta_spd(n, b, a, seed=seed, gamma=0.0)
        self.n, self.b, self.a = n, b, a
        self.U = [torch.from_numpy(x).to(device) for x in (D, L, R, T)]
        self.device = device

    def blocks(self, gamma, out=None):
        n, b, a = self.n, self.b, self.a
        if out is None:
            out = [torch.empty_like(x) for x in self.U]
        for o, u in zip(out, self.U):
            o.copy_(u)
        out[0].diagonal(dim1=1, dim2=2).add_(gamma * (3 * b + a))
        if a:
            out[3].diagonal().add_(gamma * (n * b + a))
        return out


def _is_pd(sh, gamma, work):
    import serinv_amd.algs as algs

    blocks = sh.blocks(gamma, work)
    try:
        algs.pobtaf(*blocks)
        return True
    except np.linalg.LinAlgError:
        return False


def lambda_max(sh, gamma, iters=60, seed=1):
    from serinv_amd.mixed import residual

    A = sh.blocks(gamma)
    N = sh.n * sh.b + sh.a
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.rand(N, 1, generator=g, dtype=torch.float64).to(sh.device)
    x /= x.norm()
    zero = torch.zeros_like(x)
    lam = 0.0
    for _ in range(iters):
        y = -residual(A, x, zero)  # A x
        lam = float((x * y).sum())
        x = y / y.norm()
    return lam


def lambda_min(sh, gamma, iters=60, seed=2):
    import serinv_amd.algs as algs

    F = sh.blocks(gamma)
    algs.pobtaf(*F)
    N = sh.n * sh.b + sh.a
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.randn(N, 1, generator=g, dtype=torch.float64).to(sh.device)
    x /= x.norm()
    mu = 0.0
    for _ in range(iters):
        y = x.clone()
        algs.pobtas(*F, y, trans="N")
        algs.pobtas(*F, y, trans="C")
        mu = float((x * y).sum())  # Rayleigh quotient of A^{-1}
        x = y / y.norm()
    return 1.0 / mu


def condition(sh, gamma, iters=60):
    return lambda_max(sh, gamma, iters) / lambda_min(sh, gamma, iters)


def pd_boundary(sh, lo=0.0, hi=1.0, rel=1e-10):
    work = [torch.empty_like(x) for x in sh.U]
    if _is_pd(sh, lo, work):
        return lo
    while hi - lo > rel * hi:
        mid = 0.5 * (lo + hi)
        if _is_pd(sh, mid, work):
            hi = mid
        else:
            lo = mid
    return hi


def gamma_for(sh, kappa, gamma0, steps=22, iters=60):
    """gamma with cond(A(gamma)) ~ kappa (bisection on log(gamma - gamma0))."""
    lo, hi = math.log(max(gamma0 * 1e-12, 1e-300)), math.log(max(1.0 - gamma0, 1e-300))
    best = None
    for _ in range(steps):
        mid = 0.5 * (lo + hi)
        c = condition(sh, gamma0 + math.exp(mid), iters)
        best = (gamma0 + math.exp(mid), c)
        if c > kappa:
            lo = mid
        else:
            hi = mid
        if abs(math.log(c / kappa)) < 0.05:
            break
    return best


def calibrate(n, b, a, kappas, device, seed=63, path=TABLE):
    sh = ShiftedBTA(n, b, a, device, seed)
    g0 = pd_boundary(sh)
    print(f"[calibrate] n={n} b={b} a={a}: PD boundary gamma0 = {g0:.12g}", flush=True)
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        table = {}
    for k in kappas:
        gamma, c = gamma_for(sh, k, g0)
        print(f"[calibrate] target {k:.0e}: gamma = {gamma:.12g}, cond = {c:.3e}", flush=True)
        table[f"{n},{b},{a},{k:.0e}"] = {"gamma": gamma, "cond": c, "gamma_pd": g0, "seed": seed}
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    return table


if __name__ == "__main__":
    import argparse

    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=256)
    p.add_argument("--b", type=int, default=1024)
    p.add_argument("--a", type=int, default=32)
    p.add_argument("--kappa", type=float, nargs="+", default=[1e4, 1e6, 1e8])
    p.add_argument("--out", default=TABLE)
    args = p.parse_args()
    out = calibrate(args.n, args.b, args.a, args.kappa, torch.device("cuda:0"), path=args.out)
    print(json.dumps(out, indent=1))
