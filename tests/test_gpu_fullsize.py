"""Parity at the benchmark workloads: n=256, b=1024, a=32, through the C ABI.

No dense matrix of this size (N = 262,176) fits, so the FP64 path is checked
by size-independent properties, computed with FP64 block products on the GPU
(tests/_blockprod.py) from the ORIGINAL blocks A of the calibrated matrix:

* pobtaf:  L L^T reproduces A on the BTA pattern, max |L L^T - A| / max |A|
           per block array <= 1e-12 (backward error of a Cholesky
           factorization, u * growth); the factor's upper triangles are zero.
* pobtas:  ||A x - B||_F / ||B||_F <= 1e-12 for r = 1 and r = 64 (N then C).
* pobtasi: every diagonal block and the tip block of A X is the identity:
           max |(AX)_ii - I| <= 100 * cond(A) * u (the residual bound of a
           computed inverse; cond from the calibration table).
The oracle comparison at this size (the NumPy restatement run on the host)
is part of bench.py, which reports the per-output errors.

The mixed FP32/FP64 form (cond 1e4, 1e6, 1e8) is compared with the FP64
outputs pinned above, and its FP64 iterative refinement on 16 right-hand
sides must reach the FP64 residual floor.  The batched plan runs the 64
uneven systems of the batched benchmark and every system is compared with
the oracle run on that system alone at 1e-12.
"""

import json
import os

import numpy as np
import pytest

from tests._blockprod import ax_diag, bta_matvec, identity_err, llt_pattern, maxrel

pytestmark = pytest.mark.gpu
N_, B_, A_ = 256, 1024, 32
U64 = 2.0 ** -53


def _table():
    from serinv_amd.utils.conditioning import TABLE

    with open(TABLE) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def shifted(gpu_device):
    from serinv_amd.utils.conditioning import ShiftedBTA

    sh = ShiftedBTA(N_, B_, A_, gpu_device)
    yield sh
    del sh
    import torch

    torch.cuda.empty_cache()


def _rhs(dev, r, seed):
    import torch

    rng = np.random.default_rng(seed)
    return torch.from_numpy(rng.uniform(-1.0, 1.0, (N_ * B_ + A_, r))).to(dev)


def test_fp64_full_size_kappa1e6(shifted, gpu_device):
    import torch

    import serinv_amd.algs as algs

    entry = _table()[f"{N_},{B_},{A_},{1e6:.0e}"]
    A = shifted.blocks(entry["gamma"])
    F = [x.clone() for x in A]
    algs.pobtaf(*F)  # raises on a non-positive pivot or a scheduling time-out
    P = llt_pattern(F)
    errs = {nm: maxrel(p, x) for nm, p, x in zip("DLRT", P, A)}
    del P
    print(f"\n[fullsize] cond {entry['cond']:.3e}: max|LL^T - A|/max|A| {errs}", flush=True)
    for nm, e in errs.items():
        assert e <= 1e-12, f"L L^T - A on block array {nm}: {e:.3e}"
    assert torch.all(torch.triu(F[0], 1) == 0) and torch.all(torch.triu(F[3], 1) == 0)

    for r in (1, 64):
        B = _rhs(gpu_device, r, 70 + r)
        x = B.clone()
        algs.pobtas(*F, x, trans="N")
        algs.pobtas(*F, x, trans="C")
        res = float((bta_matvec(A, x) - B).norm() / B.norm())
        print(f"[fullsize] r={r}: ||Ax - B|| / ||B|| = {res:.3e}", flush=True)
        assert res <= 1e-12, f"solve residual r={r}: {res:.3e}"
        del x, B

    algs.pobtasi(*F)
    P, PT = ax_diag(A, F)
    e_d, e_t = identity_err(P), identity_err(PT)
    bar = 100.0 * entry["cond"] * U64
    print(f"[fullsize] max|(AX)_ii - I| = {e_d:.3e}, tip {e_t:.3e} (bar {bar:.2e})", flush=True)
    assert e_d <= bar and e_t <= bar
    # X_ii and X_T are stored as full symmetric blocks
    assert maxrel(F[0], F[0].transpose(1, 2)) == 0.0 and maxrel(F[3], F[3].T) == 0.0
    del P, PT, F, A
    torch.cuda.empty_cache()


# Bars of the mixed form against the FP64 outputs (max relative error per
# block array): FP32 products carry ~cond * growth * 6e-8; measured at this
# workload 1.4e-4 (X) and 1.1e-4 (L) at cond 1e8, below 1e-5 at cond <= 1e6.
MIXED_BAR = {1e4: 1e-4, 1e6: 1e-4, 1e8: 1e-3}


@pytest.mark.parametrize("kappa", [1e4, 1e6, 1e8])
def test_mixed_full_size(shifted, gpu_device, kappa):
    import torch

    import serinv_amd.algs as algs
    from serinv_amd.logdet import pobta_logdet
    from serinv_amd.mixed import pobtaf_mixed, pobtas_refine, pobtasi_mixed

    entry = _table()[f"{N_},{B_},{A_},{kappa:.0e}"]
    A = shifted.blocks(entry["gamma"])
    F64 = [x.clone() for x in A]
    algs.pobtaf(*F64)
    F32 = [x.clone() for x in A]
    pobtaf_mixed(*F32)
    eL = {nm: maxrel(p, q) for nm, p, q in zip("DLRT", F32, F64)}
    ld64, ld32 = float(pobta_logdet(F64[0], F64[3])), float(pobta_logdet(F32[0], F32[3]))
    eld = abs(ld32 - ld64) / abs(ld64)
    # FP64 iterative refinement, 16 right-hand sides
    B = _rhs(gpu_device, 16, 80)
    x = B.clone()
    hist = pobtas_refine(A, F32, x, max_iter=10, tol=1e-14)
    res = (bta_matvec(A, x) - B)
    rel_res = float(res.norm() / B.norm())
    ones = torch.ones_like(x[:, :1])
    anorm = float(bta_matvec([t.abs() for t in A], ones).max())  # ||A||_inf
    eta = float(res.abs().max()) / (anorm * float(x.abs().max()) + float(B.abs().max()))
    del res, x, B
    X64 = F64
    algs.pobtasi(*X64)
    X32 = F32
    pobtasi_mixed(*X32)
    eX = {nm: maxrel(p, q) for nm, p, q in zip("DLRT", X32, X64)}
    print(f"\n[fullsize mixed] cond {entry['cond']:.2e}: L {eL} X {eX} logdet {eld:.2e} "
          f"refine {hist} rel. residual {rel_res:.2e} backward error {eta:.2e}", flush=True)
    for nm in "DLRT":
        assert eL[nm] <= MIXED_BAR[kappa], f"mixed factor {nm}: {eL[nm]:.3e}"
        assert eX[nm] <= MIXED_BAR[kappa], f"mixed selected inverse {nm}: {eX[nm]:.3e}"
    assert eld <= 1e-9
    # refinement reaches the FP64 floor: the normwise backward error
    # ||B - A x||_max / (||A||_inf ||x||_max + ||B||_max) is at the level of
    # the FP64 rounding of the residual itself at every cond; the relative
    # residual reaches 1e-12 where cond * u allows it
    assert eta <= 1e-14, eta
    if kappa <= 1e6:
        assert rel_res <= 1e-12, rel_res
    del A, F64, F32
    torch.cuda.empty_cache()


def _bench_shapes(nmat=64):
    rng = np.random.default_rng(63)
    return [(int(rng.integers(16, 257)), int(rng.choice([64, 128, 256])), int(rng.integers(4, 33)))
            for _ in range(nmat)]


def test_batched_64_workload(gpu_device):
    """The batched benchmark's 64 systems (seed 63 + k, 4 RHS), one graph
    replay, every output against the oracle run on that system alone (one
    BLAS thread: at b <= 256 threaded OpenBLAS calls cost more than they
    save)."""
    from threadpoolctl import threadpool_limits

    from oracle import serinv_np
    from serinv_amd.batched import BatchedBTA
    from serinv_amd.utils.matrices import bta_spd
    from tests._helpers import relerr, to_dev, to_host

    shapes = _bench_shapes()
    host, dev = [], []
    for k, (n, b, a) in enumerate(shapes):
        arrs = bta_spd(n, b, a, seed=63 + k, n_rhs=4)
        host.append(arrs)
        dev.append(to_dev(gpu_device, *arrs))
    plan = BatchedBTA([d[:4] for d in dev], rhs=[d[4] for d in dev], selinv=True, graph=True)
    plan.run()
    info, logdet, status = plan.results()
    plan.close()
    assert status == 0 and np.all(info == 0), (info, status)
    worst = 0.0
    with threadpool_limits(limits=1):
        for k, (h, d) in enumerate(zip(host, dev)):
            D, L, R, T, B = (x.copy() for x in h)
            serinv_np.pobtaf(D, L, R, T)
            ld = serinv_np.logdet_from_factor(D, T)
            serinv_np.pobtas(D, L, R, T, B, trans="N")
            serinv_np.pobtas(D, L, R, T, B, trans="C")
            serinv_np.pobtasi(D, L, R, T)
            errs = [abs(logdet[k] - ld) / abs(ld), relerr(to_host(d[4])[0], B)]
            errs += [relerr(g, o) for g, o in zip(to_host(*d[:4]), (D, L, R, T))]
            worst = max(worst, max(errs))
            assert max(errs) < 1e-12, f"system {k} {shapes[k]}: {errs}"
            if k % 16 == 15:
                print(f"\n[fullsize batched] systems 0..{k} checked, worst {worst:.2e}", flush=True)
    print(f"\n[fullsize batched] 64 systems, worst rel. error vs oracle {worst:.2e}", flush=True)
