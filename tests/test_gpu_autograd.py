"""GPU checks of the differentiable log-determinant / quadratic form
(serinv_amd.autograd.pobta_logdet_quad).

There is no reference counterpart; the checks pin it by mathematics:
* forward values against numpy.linalg.slogdet and a dense solve;
* block gradients against the dense formulas dlogdet/dA = A^{-1} and
  dquad/dA = -x x^T (x = A^{-1} y) mapped to the stored entries;
* theta-gradients of f(theta) = logdet A(theta) - y^T A(theta)^{-1} y with
  A(theta) = theta1 K + theta2 M + theta3 I, i.e.
  df/dtheta = tr(A^{-1} dA/dtheta) + x^T (dA/dtheta) x, against dense FP64
  inversion and central finite differences (h = 1e-6).
Tolerances: 1e-10 relative against dense FP64, 1e-6 against finite
differences (their truncation and rounding error is about 1e-9 here).
"""

import numpy as np
import pytest

from oracle import dense
from serinv_amd.utils.matrices import bta_spd
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu


def _weights(m):
    w = np.tril(np.full((m, m), 2.0), -1)
    np.fill_diagonal(w, 1.0)
    return w


@pytest.mark.parametrize("n,b,a,r", [(4, 16, 4, 2), (3, 70, 5, 1), (2, 64, 0, 3), (5, 8, 65, 1)])
def test_block_gradients_dense(gpu_device, n, b, a, r):
    import torch

    from serinv_amd.autograd import pobta_logdet_quad

    D, L, R, T, y = bta_spd(n, b, a, seed=70 + n + b, n_rhs=r)
    A = dense.bta_arrays_to_dense(D, L, R, T)
    Ainv = np.linalg.inv(A)
    x = Ainv @ y
    d, l, rr, t, yy = (v.requires_grad_(True) for v in to_dev(gpu_device, D, L, R, T, y))
    ld, q = pobta_logdet_quad(d, l, rr, t, yy)
    assert abs(float(ld.detach()) - np.linalg.slogdet(A)[1]) < 1e-12 * abs(np.linalg.slogdet(A)[1])
    assert abs(float(q.detach()) - float(np.sum(y * x))) < 1e-12 * abs(float(np.sum(y * x)))
    gl, gq = 0.7, -1.3
    (gl * ld + gq * q).backward()
    G = gl * Ainv - gq * (x @ x.T)  # d f / dA (dense, symmetric)
    Gd, Gl, Gr, Gt = dense.bta_dense_to_arrays(G, b, a, n)
    wD = _weights(b)
    refD = Gd * wD
    refL = 2.0 * Gl
    refR = 2.0 * Gr
    refT = Gt * _weights(a) if a else Gt
    for nm, g, ref in (("D", d.grad, refD), ("L", l.grad, refL), ("R", rr.grad, refR), ("T", t.grad, refT)):
        assert relerr(to_host(g)[0], ref) < 1e-10, f"grad {nm}: {relerr(to_host(g)[0], ref):.3e}"
    assert relerr(to_host(yy.grad)[0], 2.0 * gq * x) < 1e-10


def _theta_problem(n, b, a, dev, seed=63):
    import torch

    K = bta_spd(n, b, a, seed=seed)[:4]
    M = bta_spd(n, b, a, seed=seed + 1)[:4]
    rng = np.random.default_rng(seed + 3)
    y = rng.standard_normal(n * b + a)
    Kt = to_dev(dev, *K)
    Mt = to_dev(dev, *M)
    yt = torch.from_numpy(y).to(dev)

    def blocks(theta):
        eyeb = torch.eye(b, dtype=torch.float64, device=dev)
        eyea = torch.eye(a, dtype=torch.float64, device=dev)
        Dm = theta[0] * Kt[0] + theta[1] * Mt[0] + theta[2] * eyeb
        Lm = theta[0] * Kt[1] + theta[1] * Mt[1]
        Rm = theta[0] * Kt[2] + theta[1] * Mt[2]
        Tm = theta[0] * Kt[3] + theta[1] * Mt[3] + theta[2] * eyea
        return Dm.contiguous(), Lm.contiguous(), Rm.contiguous(), Tm.contiguous()

    return K, M, y, yt, blocks


def _f(theta, blocks, yt):
    from serinv_amd.autograd import pobta_logdet_quad

    ld, q = pobta_logdet_quad(*blocks(theta), yt)
    return ld - q


@pytest.mark.parametrize("n,b,a", [(4, 16, 4), (6, 96, 8)])
def test_theta_gradient_dense_and_fd(gpu_device, n, b, a):
    import torch

    K, M, y, yt, blocks = _theta_problem(n, b, a, gpu_device)
    theta0 = np.array([1.3, 0.6, 1.7])
    theta = torch.tensor(theta0, dtype=torch.float64, device=gpu_device, requires_grad=True)
    f = _f(theta, blocks, yt)
    f.backward()
    g = theta.grad.cpu().numpy()
    # dense FP64
    Kd, Md = dense.bta_arrays_to_dense(*K), dense.bta_arrays_to_dense(*M)
    A = theta0[0] * Kd + theta0[1] * Md + theta0[2] * np.eye(Kd.shape[0])
    Ainv = np.linalg.inv(A)
    x = Ainv @ y
    # f = logdet - y^T A^{-1} y  =>  df/dtheta = tr(A^{-1} dA) + x^T dA x
    ref = np.array([np.sum(Ainv * dA) + x @ dA @ x for dA in (Kd, Md, np.eye(Kd.shape[0]))])
    assert abs(float(f.detach()) - (np.linalg.slogdet(A)[1] - y @ x)) < 1e-11 * abs(float(f.detach()))
    assert np.max(np.abs(g - ref)) < 1e-10 * np.max(np.abs(ref)), (g, ref)
    # central finite differences
    h = 1e-6
    fd = np.zeros(3)
    with torch.no_grad():
        for k in range(3):
            tp, tm = theta0.copy(), theta0.copy()
            tp[k] += h
            tm[k] -= h
            fp = float(_f(torch.tensor(tp, device=gpu_device), blocks, yt))
            fm = float(_f(torch.tensor(tm, device=gpu_device), blocks, yt))
            fd[k] = (fp - fm) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-6 * np.max(np.abs(fd)), (g, fd)


def _tri_inv(C, base=1024):
    """Inverse of a lower-triangular dense GPU matrix by 2 x 2 block recursion
    (matrix products only above `base`; test-only)."""
    import torch

    m = C.shape[0]
    if m <= base:
        return torch.linalg.solve_triangular(
            C, torch.eye(m, dtype=C.dtype, device=C.device), upper=False)
    h = m // 2
    A11 = _tri_inv(C[:h, :h].contiguous(), base)
    A22 = _tri_inv(C[h:, h:].contiguous(), base)
    out = torch.zeros_like(C)
    out[:h, :h] = A11
    out[h:, h:] = A22
    out[h:, :h] = -(A22 @ (C[h:, :h] @ A11))
    return out


def test_theta_gradient_workload_n128_b256_a8(gpu_device):
    """The differentiable workload: K, M seeded SPD BTA (n=128, b=256, a=8;
    seeds 63, 64), y ~ N(0, 1); gradient against dense FP64 (torch on the GPU,
    test-only) and central finite differences."""
    import torch

    n, b, a = 128, 256, 8
    K, M, y, yt, blocks = _theta_problem(n, b, a, gpu_device)
    theta0 = np.array([0.8, 1.4, 1.1])
    theta = torch.tensor(theta0, dtype=torch.float64, device=gpu_device, requires_grad=True)
    f = _f(theta, blocks, yt)
    f.backward()
    g = theta.grad.cpu().numpy()
    h = 1e-6
    fd = np.zeros(3)
    with torch.no_grad():
        for k in range(3):
            tp, tm = theta0.copy(), theta0.copy()
            tp[k] += h
            tm[k] -= h
            fd[k] = (float(_f(torch.tensor(tp, device=gpu_device), blocks, yt)) -
                     float(_f(torch.tensor(tm, device=gpu_device), blocks, yt))) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-6 * np.max(np.abs(fd)), (g, fd)
    # dense FP64 on the GPU (torch linalg; test-only check, 8.6 GB per dense
    # matrix): A^{-1} = C^{-T} C^{-1} from the dense Cholesky factor C
    with torch.no_grad():
        Db, Lb, Rb, Tb = (t.cpu().numpy() for t in blocks(torch.tensor(theta0, device=gpu_device)))
        C = torch.linalg.cholesky(torch.from_numpy(dense.bta_arrays_to_dense(Db, Lb, Rb, Tb)).to(gpu_device))
        Kd = torch.from_numpy(dense.bta_arrays_to_dense(*K)).to(gpu_device)
        Md = torch.from_numpy(dense.bta_arrays_to_dense(*M)).to(gpu_device)
        Ci = _tri_inv(C)
        del C
        Ainv = Ci.mT @ Ci
        del Ci
        x = Ainv @ yt
        tr = [float((Ainv * Kd).sum()), float((Ainv * Md).sum()), float(torch.trace(Ainv))]
        del Ainv
        ref = np.array([tr[0] + float(x @ (Kd @ x)), tr[1] + float(x @ (Md @ x)), tr[2] + float(x @ x)])
        del Kd, Md
    assert np.max(np.abs(g - ref)) < 1e-10 * np.max(np.abs(ref)), (g, ref)


def test_backward_runs_once(gpu_device):
    import torch

    from serinv_amd.autograd import pobta_logdet_quad

    D, L, R, T, y = bta_spd(2, 8, 2, n_rhs=1)
    d, l, r, t, yy = (v.requires_grad_(True) for v in to_dev(gpu_device, D, L, R, T, y))
    ld, q = pobta_logdet_quad(d, l, r, t, yy)
    (ld + q).backward(retain_graph=True)
    with pytest.raises(RuntimeError):
        (ld + q).backward()
