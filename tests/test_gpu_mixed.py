"""GPU checks of the mixed FP32/FP64 form (serinv_amd.mixed).

* rocbta_dresidual against a dense FP64 B - A X (1e-13 relative), reading
  only the lower triangles of D and T (their upper triangles are filled with
  garbage on the device);
* the mixed factor, log-determinant and selected inverse against the FP64
  oracle on diagonally dominant matrices: their errors are those of FP32
  products, so the bars are 1e-5 (factor, selected inverse) and 1e-7
  (log-determinant, whose sum is FP64);
* FP64 iterative refinement (pobtas_refine) with the mixed factor reaching
  the FP64 parity bar 1e-12 against the oracle solve.
Both pobtaf algorithms (persistent chain, launch-per-tile) run every case.
"""

import numpy as np
import pytest

from oracle import dense, serinv_np
from serinv_amd.utils.matrices import bta_spd
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["chain", "loop"])
def algo(request, monkeypatch):
    if request.param == "loop":
        monkeypatch.setenv("ROCBTA_POBTAF_ALGO", "loop")
    else:
        monkeypatch.delenv("ROCBTA_POBTAF_ALGO", raising=False)
    return request.param


@pytest.mark.parametrize("n,b,a,r", [(3, 64, 16, 1), (2, 70, 0, 3), (4, 100, 70, 9), (1, 8, 5, 16),
                                     (5, 128, 64, 4)])
def test_residual_dense(gpu_device, n, b, a, r):
    from serinv_amd.mixed import residual

    D, L, R, T, B = bta_spd(n, b, a, seed=40 + n + b, n_rhs=r)
    rng = np.random.default_rng(7)
    X = rng.standard_normal(B.shape)
    A = dense.bta_arrays_to_dense(D, L, R, T)
    ref = B - A @ X
    Dg, Tg = D.copy(), T.copy()
    iu = np.triu_indices(b, 1)
    Dg[:, iu[0], iu[1]] = 1e6  # only the lower triangles may be read
    if a:
        Tg[np.triu_indices(a, 1)] = -1e6
    d, l, rr, t, x, bb = to_dev(gpu_device, Dg, L, R, Tg, X, B)
    out = residual((d, l, rr, t), x, bb)
    scale = np.max(np.abs(B)) + np.max(np.abs(A)) * np.max(np.abs(X)) * (3 * b + a)
    assert np.max(np.abs(to_host(out)[0] - ref)) < 1e-13 * scale


@pytest.mark.parametrize("n,b,a", [(4, 64, 16), (3, 130, 33), (2, 256, 8), (5, 70, 0)])
def test_mixed_factor_selinv(gpu_device, algo, n, b, a):
    from serinv_amd.logdet import pobta_logdet
    from serinv_amd.mixed import pobtaf_mixed, pobtasi_mixed

    D, L, R, T, _ = bta_spd(n, b, a, seed=90 + n + b)
    Do, Lo, Ro, To = (x.copy() for x in (D, L, R, T))
    serinv_np.pobtaf(Do, Lo, Ro, To)
    ld_ref = serinv_np.logdet_from_factor(Do, To)
    d, l, r, t = to_dev(gpu_device, D, L, R, T)
    pobtaf_mixed(d, l, r, t)
    for nm, g, o in zip("DLRT", to_host(d, l, r, t), (Do, Lo, Ro, To)):
        assert relerr(g, o) < 1e-5, f"factor {nm}: {relerr(g, o):.2e}"
    # the FP32 products are visible: the mixed factor differs from the FP64 one
    assert relerr(to_host(l)[0], Lo) > 1e-13 or n == 1
    ld = float(pobta_logdet(d, t))
    assert abs(ld - ld_ref) < 1e-7 * abs(ld_ref)
    serinv_np.pobtasi(Do, Lo, Ro, To)
    pobtasi_mixed(d, l, r, t)
    for nm, g, o in zip("DLRT", to_host(d, l, r, t), (Do, Lo, Ro, To)):
        assert relerr(g, o) < 1e-5, f"selected inverse {nm}: {relerr(g, o):.2e}"


@pytest.mark.parametrize("n,b,a", [(4, 64, 16), (3, 130, 33), (2, 256, 8)])
def test_mixed_factor_open_last_block(gpu_device, algo, n, b, a):
    """pobtaf_mixed(factorize_last_block=False): blocks 0 .. n-2 as the FP64
    oracle's factor, the last block, its arrow block and the tip as the
    oracle's Schur-updated values (pobtaf.py:143-159), at the mixed bar."""
    from serinv_amd.mixed import pobtaf_mixed

    D, L, R, T, _ = bta_spd(n, b, a, seed=95 + n + b)
    Do, Lo, Ro, To = (x.copy() for x in (D, L, R, T))
    serinv_np.pobtaf(Do, Lo, Ro, To, factorize_last_block=False)
    d, l, r, t = to_dev(gpu_device, D, L, R, T)
    pobtaf_mixed(d, l, r, t, factorize_last_block=False)
    for nm, g, o in zip("DLRT", to_host(d, l, r, t), (Do, Lo, Ro, To)):
        assert relerr(g, o) < 1e-5, f"{nm}: {relerr(g, o):.2e}"


@pytest.mark.parametrize("n,b,a,r", [(4, 64, 16, 1), (3, 130, 33, 16), (6, 96, 0, 4)])
def test_refinement_reaches_fp64(gpu_device, algo, n, b, a, r):
    from serinv_amd.mixed import pobtaf_mixed, pobtas_refine

    D, L, R, T, B = bta_spd(n, b, a, seed=11 + n, n_rhs=r, gamma=0.4)
    A = to_dev(gpu_device, D, L, R, T)
    F = [x.clone() for x in A]
    pobtaf_mixed(*F)
    bb = to_dev(gpu_device, B)[0]
    hist = pobtas_refine(A, F, bb, max_iter=10, tol=1e-14)
    xo = np.linalg.solve(dense.bta_arrays_to_dense(D, L, R, T), B)
    assert hist[0] > 1e-12  # the plain mixed solve is not FP64-accurate
    assert hist[-1] < 1e-13, hist
    assert relerr(to_host(bb)[0], xo) < 1e-12
