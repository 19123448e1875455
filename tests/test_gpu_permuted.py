"""GPU parity of the permuted ("middle partition") routines.

rocbta_dpobtaf_permuted / _dpobtasi_permuted / _dpobtas_permuted (through
serinv_amd.algs with buffer=) against the oracle's restatement of
src/serinv/algs/pobtaf.py:186-281, pobtasi.py:175-244, pobtas.py:164-218
(and the BT forms pobtf.py:125-185, pobtsi.py:128-179, pobts.py:121-165),
and through serinv's reduced-system procedure (tests/_permuted.py) against
the dense inverse and dense solve, as the reference's permuted tests do.
Tolerance: 1e-12 relative (max-norm per output array) for the factor and the
reduced system against the oracle, 1e-11 for the completed inverse and
solve against dense LAPACK (the procedure chains three routines).
Shapes: the reference's grid (n 3, 4; b, a in {2, 3}), n = 2 (no middle
block), the arrowless case, 64-wide tile boundaries (b 63, 65, 100, 130,
a 1, 64, 65) and n = 16, b = 256, a = 64.
"""

import numpy as np
import pytest

from oracle import dense, generators, serinv_np
from tests._helpers import relerr, to_dev, to_host
from tests._permuted import selinv_permuted, solve_permuted

pytestmark = pytest.mark.gpu
TOL = 1e-12
TOL_DENSE = 1e-11


@pytest.fixture(scope="module")
def algs(gpu_device):
    import serinv_amd.algs as A

    return A


def case(b, a, n, seed):
    rng = np.random.default_rng(seed)
    A = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(A)
    return A


SHAPES = [(2, 2, 3), (3, 3, 4), (2, 3, 2), (3, 0, 4), (63, 1, 3), (65, 64, 5), (100, 65, 4),
          (130, 16, 6), (64, 0, 7), (256, 64, 16)]


class GpuRT:
    """serinv_amd.algs under the reference's call signatures."""

    def __init__(self, algs):
        self.pobtaf = algs.pobtaf
        self.pobtasi = algs.pobtasi
        self.pobtas = algs.pobtas


def gpu_zeros(dev):
    import torch

    return lambda shape: torch.zeros(shape, dtype=torch.float64, device=dev)


@pytest.mark.parametrize("b,a,n", SHAPES)
def test_pobtaf_permuted_matches_oracle(algs, gpu_device, b, a, n):
    import torch

    print(f"pobtaf_permuted b={b} a={a} n={n}", flush=True)
    A = case(b, a, n, seed=11000 + b + 7 * a + 13 * n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    gD, gL, gR, gT = to_dev(gpu_device, D, L, R, T)
    gbuf = torch.full_like(gD, float("nan"))  # buffer[0] is never read
    algs.pobtaf(gD, gL, gR, gT, buffer=gbuf)
    serinv_np.pobtaf(D, L, R, T, buffer=buf)
    hD, hL, hR, hT, hbuf = to_host(gD, gL, gR, gT, gbuf)
    for x, y in ((hD, D), (hL, L), (hR, R), (hT, T), (hbuf[1:], buf[1:])):
        assert relerr(x, y) < TOL
    assert np.all(np.triu(hD[1:n - 1], 1) == 0)


@pytest.mark.parametrize("b,a,n", SHAPES)
def test_pobtasi_permuted_matches_oracle_and_dense(algs, gpu_device, b, a, n):
    print(f"pobtasi_permuted b={b} a={a} n={n}", flush=True)
    A = case(b, a, n, seed=12000 + b + 7 * a + 13 * n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    gD, gL, gR, gT, gbuf = to_dev(gpu_device, D, L, R, T, np.zeros_like(D))
    selinv_permuted(gD, gL, gR, gT, gbuf, GpuRT(algs), gpu_zeros(gpu_device))
    # the oracle through the same procedure
    oD, oL, oR, oT, obuf = D.copy(), L.copy(), R.copy(), T.copy(), np.zeros_like(D)

    class O:
        pobtaf = staticmethod(serinv_np.pobtaf)
        pobtasi = staticmethod(serinv_np.pobtasi)

    selinv_permuted(oD, oL, oR, oT, obuf, O, np.zeros)
    hD, hL, hR, hT, hbuf = to_host(gD, gL, gR, gT, gbuf)
    for x, y in ((hD, oD), (hL, oL), (hR, oR), (hT, oT), (hbuf[1:], obuf[1:])):
        assert relerr(x, y) < TOL_DENSE
    X = np.linalg.inv(A)
    Dr, Lr, Rr, Tr = dense.bta_dense_to_arrays(X, b, a, n)
    for x, y in ((hD, Dr), (hL, Lr), (hR, Rr), (hT, Tr)):
        assert relerr(x, y) < TOL_DENSE
    for i in range(1, n):
        assert relerr(hbuf[i], X[:b, i * b:(i + 1) * b]) < TOL_DENSE


@pytest.mark.parametrize("r", (1, 3, 16))
@pytest.mark.parametrize("b,a,n", SHAPES)
def test_pobtas_permuted_matches_dense_solve(algs, gpu_device, b, a, n, r):
    print(f"pobtas_permuted b={b} a={a} n={n} r={r}", flush=True)
    A = case(b, a, n, seed=13000 + b + 7 * a + 13 * n + r)
    Bm = np.random.default_rng(r).uniform(-1, 1, (n * b + a, r))
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    gD, gL, gR, gT, gbuf, gB = to_dev(gpu_device, D, L, R, T, np.zeros_like(D), Bm)
    solve_permuted(gD, gL, gR, gT, gbuf, gB, GpuRT(algs), gpu_zeros(gpu_device))
    (hB,) = to_host(gB)
    assert relerr(hB, np.linalg.solve(A, Bm)) < TOL_DENSE


@pytest.mark.parametrize("trans", ("N", "C"))
@pytest.mark.parametrize("b,a,n", [(65, 64, 5), (100, 3, 4), (64, 0, 6)])
def test_pobtas_permuted_half_matches_oracle(algs, gpu_device, b, a, n, trans):
    """One half on its own, from the oracle's factor, against the oracle."""
    print(f"pobtas_permuted half {trans} b={b} a={a} n={n}", flush=True)
    A = case(b, a, n, seed=14000 + b + a + n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    serinv_np.pobtaf(D, L, R, T, buffer=buf)
    Bm = np.random.default_rng(2).uniform(-1, 1, (n * b + a, 5))
    gD, gL, gR, gT, gbuf, gB = to_dev(gpu_device, D, L, R, T, buf, Bm)
    algs.pobtas(gD, gL, gR, gT, gB, buffer=gbuf, trans=trans)
    B = Bm.copy()
    serinv_np.pobtas(D, L, R, T, B, buffer=buf, trans=trans)
    (hB,) = to_host(gB)
    assert relerr(hB, B) < TOL


@pytest.mark.parametrize("b,n", [(3, 4), (64, 5), (130, 3)])
def test_bt_permuted_routines(algs, gpu_device, b, n):
    """pobtf / pobtsi / pobts with buffer= (a = 0 through the same entry
    points) against the dense inverse and solve."""
    print(f"bt permuted b={b} n={n}", flush=True)
    rng = np.random.default_rng(15000 + b + n)
    A = generators.dd_bt(b, n, rng)
    generators.symmetrize(A)

    class BT:
        @staticmethod
        def pobtaf(D, L, R, T, **kw):
            algs.pobtf(D, L, **kw)

        @staticmethod
        def pobtasi(D, L, R, T, **kw):
            algs.pobtsi(D, L, **kw)

        @staticmethod
        def pobtas(D, L, R, T, B, **kw):
            algs.pobts(D, L, B, **kw)

    import torch

    D, L = dense.bt_dense_to_arrays(A, b, n)
    R = torch.zeros((n, 0, b), dtype=torch.float64, device=gpu_device)
    T = torch.zeros((0, 0), dtype=torch.float64, device=gpu_device)
    gD, gL, gbuf = to_dev(gpu_device, D, L, np.zeros_like(D))
    selinv_permuted(gD, gL, R, T, gbuf, BT, gpu_zeros(gpu_device))
    Dr, Lr = dense.bt_dense_to_arrays(np.linalg.inv(A), b, n)
    hD, hL = to_host(gD, gL)
    assert relerr(hD, Dr) < TOL_DENSE and relerr(hL, Lr) < TOL_DENSE
    Bm = rng.uniform(-1, 1, (n * b, 3))
    gD, gL, gbuf, gB = to_dev(gpu_device, D, L, np.zeros_like(D), Bm)
    solve_permuted(gD, gL, R, T, gbuf, gB, BT, gpu_zeros(gpu_device))
    (hB,) = to_host(gB)
    assert relerr(hB, np.linalg.solve(A, Bm)) < TOL_DENSE


def test_permuted_argument_errors(algs, gpu_device):
    import torch

    from serinv_amd.utils import allocate_pobtax_permutation_buffers

    b, a, n = 4, 2, 3
    D, L, R, T = to_dev(gpu_device, *dense.bta_dense_to_arrays(case(b, a, n, 1), b, a, n))
    buf = allocate_pobtax_permutation_buffers(D)
    assert buf.shape == D.shape and buf.device == D.device and buf.dtype == D.dtype
    with pytest.raises(ValueError, match="buffer must have the shape"):
        algs.pobtaf(D, L, R, T, buffer=buf[:2])
    with pytest.raises(NotImplementedError):
        algs.pobtaf(D, L, R, T, buffer=buf, device_streaming=True)
    B = torch.zeros((n * b + a, 1), dtype=torch.float64, device=gpu_device)
    with pytest.raises(NotImplementedError, match="permuted arrowhead"):
        algs.pobtas(D, L, R, T, B, buffer=buf, device_streaming=True)
    with pytest.raises(NotImplementedError, match="Permuted arrowhead"):
        algs.pobts(D, L, B[:n * b], buffer=buf, device_streaming=True)
    with pytest.raises(ValueError, match="Invalid transpose"):
        algs.pobtas(D, L, R, T, B, buffer=buf, trans="X")
    D1, R1 = D[:1].clone(), R[:1].clone()
    L0 = L[:0].clone()
    with pytest.raises(IndexError):
        algs.pobtaf(D1, L0, R1, T, buffer=buf[:1].clone())
    with pytest.raises(IndexError):
        algs.pobtasi(D1, L0, R1, T, buffer=buf[:1].clone())
    with pytest.raises(NotImplementedError):
        allocate_pobtax_permutation_buffers(D, device_streaming=True)
