"""GPU parity of the wide block sweep of pobtas (64 right-hand-side columns
per launch, MFMA tiles; serinv_amd/csrc/solve_wide.hpp).

pobtas (src/serinv/algs/pobtas.py:80-159) through serinv_amd.algs with
ROCBTA_POBTAS_ALGO=wide against the oracle, both directions, at 1e-12
relative (max-norm).  Shapes: b a multiple of 16 from 64 on, several
64-column groups and ragged last groups (r 9, 17, 64, 70, 130), n = 1,
a = 0 and a = 64; and the default selection at b = 256, r = 64.
"""

import numpy as np
import pytest

from oracle import serinv_np
from serinv_amd.utils.matrices import bta_spd
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def algs(gpu_device):
    import serinv_amd.algs as A

    return A


@pytest.mark.parametrize("n,b,a,r", [(3, 64, 5, 9), (4, 80, 0, 70), (2, 128, 64, 130),
                                     (5, 256, 32, 64), (1, 64, 3, 17), (6, 96, 16, 65)])
def test_wide_solve_matches_oracle(algs, gpu_device, monkeypatch, n, b, a, r):
    monkeypatch.setenv("ROCBTA_POBTAS_ALGO", "wide")
    print(f"wide pobtas n={n} b={b} a={a} r={r}", flush=True)
    D, L, R, T, B = bta_spd(n, b, a, seed=1200 + n + b + r, n_rhs=r)
    serinv_np.pobtaf(D, L, R, T)
    d, l, rr, t, bb = to_dev(gpu_device, D, L, R, T, B)
    Bo = B.copy()
    for trans in ("N", "C"):
        serinv_np.pobtas(D, L, R, T, Bo, trans=trans)
        algs.pobtas(d, l, rr, t, bb, trans=trans)
        assert relerr(to_host(bb)[0], Bo) < TOL, trans
    # the factor is read, not written
    for g, o in zip(to_host(d, l, rr, t), (D, L, R, T)):
        assert np.array_equal(g, o)


def test_wide_solve_default_selection(algs, gpu_device, monkeypatch):
    """b >= 256 and r > 8 select the wide sweep by default; a solve of A x = B
    through it has the dense solve's answer."""
    monkeypatch.delenv("ROCBTA_POBTAS_ALGO", raising=False)
    from oracle import dense

    n, b, a, r = 3, 256, 8, 40
    D, L, R, T, B = bta_spd(n, b, a, seed=1300, n_rhs=r)
    A = dense.bta_arrays_to_dense(D, L, R, T)
    serinv_np.pobtaf(D, L, R, T)
    d, l, rr, t, bb = to_dev(gpu_device, D, L, R, T, B)
    algs.pobtas(d, l, rr, t, bb, trans="N")
    algs.pobtas(d, l, rr, t, bb, trans="C")
    assert relerr(to_host(bb)[0], np.linalg.solve(A, B)) < 1e-11
