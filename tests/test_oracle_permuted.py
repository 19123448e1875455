"""Pin the oracle's permuted ("middle partition") routines (CPU only).

serinv has no golden vectors for these routines either; its tests complete
them with the natural routines on the reduced system (block 0, block n-1,
tip) and compare with a dense inverse / dense solve
(tests/tests_algs/permuted/test_bta/test_pobtasi_permuted.py,
test_pobtas_permuted.py, and the BT versions under test_bt/), over n in
{3, 4}, b in {2, 3}, a in {2, 3}.  These tests run the same procedure
(tests/_permuted.py) over that grid plus n = 2, 5 and a = 0, and pin each
half on its own against the Cholesky of the block-permuted dense matrix:
the factorization's output block by block, the forward elimination's
output, and the backward solve given the reduced system's solution.
test_oracle_mutations.py plants mistakes that these checks catch.
"""

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import dense, generators, serinv_np
from tests._permuted import (dense_permuted_factor, permutation, selinv_permuted,
                             solve_permuted)

RTOL = 1e-12


def relerr(x, ref):
    if ref.size == 0:
        return 0.0
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def case(b, a, n, seed):
    rng = np.random.default_rng(seed)
    A = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(A)
    return A


class OracleRT:
    """The oracle under the reference's call signatures."""
    pobtaf = staticmethod(serinv_np.pobtaf)
    pobtasi = staticmethod(serinv_np.pobtasi)
    pobtas = staticmethod(serinv_np.pobtas)


def zeros(shape):
    return np.zeros(shape)


GRID = [(b, a, n) for b in (2, 3) for a in (0, 2, 3) for n in (2, 3, 4, 5)]


@pytest.mark.parametrize("b,a,n", GRID)
def test_pobtaf_permuted_vs_dense_permuted_cholesky(b, a, n):
    A = case(b, a, n, seed=5000 + 100 * b + 10 * a + n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    serinv_np.pobtaf(D, L, R, T, buffer=buf)
    ref = dense_permuted_factor(A, n, b, a)
    assert relerr(D, ref["D"]) < RTOL
    assert relerr(L, ref["L"]) < RTOL
    assert relerr(R, ref["R"]) < RTOL
    assert relerr(T, ref["T"]) < RTOL
    assert relerr(buf[1:], ref["buffer"][1:]) < RTOL
    # middle factor blocks: upper triangles zero-filled
    assert np.all(np.triu(D[1:n - 1], 1) == 0)


@pytest.mark.parametrize("b,a,n", GRID)
def test_pobtasi_permuted_vs_dense_inverse(b, a, n):
    A = case(b, a, n, seed=6000 + 100 * b + 10 * a + n)
    X = np.linalg.inv(A)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    selinv_permuted(D, L, R, T, buf, OracleRT, zeros)
    Dr, Lr, Rr, Tr = dense.bta_dense_to_arrays(X, b, a, n)
    for x, y in ((D, Dr), (L, Lr), (R, Rr), (T, Tr)):
        assert relerr(x, y) < RTOL
    # buffer[i] = X_{0,i}
    for i in range(1, n):
        assert relerr(buf[i], X[:b, i * b:(i + 1) * b]) < RTOL


@pytest.mark.parametrize("r", (1, 2, 3))
@pytest.mark.parametrize("b,a,n", GRID)
def test_pobtas_permuted_vs_dense_solve(b, a, n, r):
    A = case(b, a, n, seed=7000 + 100 * b + 10 * a + n + 7 * r)
    Bm = np.random.default_rng(r).uniform(-1, 1, (n * b + a, r))
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    B = Bm.copy()
    solve_permuted(D, L, R, T, buf, B, OracleRT, zeros)
    assert relerr(B, np.linalg.solve(A, Bm)) < RTOL


@pytest.mark.parametrize("b,a,n", [(3, 2, 5), (2, 3, 4), (4, 0, 6)])
def test_permuted_solve_halves_vs_dense(b, a, n):
    """Forward half: middle rows = L11^{-1} (P B)_mid, rest rows = (P B)_rest
    - L21 y.  Backward half, from any rest rows x_r: middle rows =
    L11^{-T} (y - L21^T x_r)."""
    A = case(b, a, n, seed=8000 + 10 * n + a)
    rng = np.random.default_rng(3)
    Bm = rng.uniform(-1, 1, (n * b + a, 2))
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    serinv_np.pobtaf(D, L, R, T, buffer=buf)
    p = permutation(n, b, a)
    m = (n - 2) * b
    Ap = A[np.ix_(p, p)]
    L11 = np.linalg.cholesky(Ap[:m, :m])
    L21 = np.linalg.solve(L11, Ap[m:, :m].T).T
    B = Bm.copy()
    serinv_np.pobtas(D, L, R, T, B, buffer=buf, trans="N")
    y = sla.solve_triangular(L11, Bm[p][:m], lower=True)
    Bp = B[p]
    assert relerr(Bp[:m], y) < RTOL
    assert relerr(Bp[m:], Bm[p][m:] - L21 @ y) < RTOL
    xr = rng.uniform(-1, 1, (2 * b + a, 2))
    B[p[m:]] = xr
    serinv_np.pobtas(D, L, R, T, B, buffer=buf, trans="C")
    xm = sla.solve_triangular(L11, y - L21.T @ xr, lower=True, trans="T")
    assert relerr(B[p[:m]], xm) < RTOL
    assert np.array_equal(B[p[m:]], xr)  # the reduced rows are not touched


@pytest.mark.parametrize("b,n", [(2, 3), (3, 4), (3, 5)])
def test_bt_permuted_routines_vs_dense(b, n):
    """pobtf / pobtsi / pobts with buffer= (src/serinv/algs/pobtf.py:125-185,
    pobtsi.py:128-179, pobts.py:121-165) through the BT reduced system."""
    rng = np.random.default_rng(9000 + 10 * b + n)
    A = generators.dd_bt(b, n, rng)
    generators.symmetrize(A)
    X = np.linalg.inv(A)

    class BT:
        @staticmethod
        def pobtaf(D, L, R, T, **kw):
            serinv_np.pobtf(D, L, **kw)

        @staticmethod
        def pobtasi(D, L, R, T, **kw):
            serinv_np.pobtsi(D, L, **kw)

        @staticmethod
        def pobtas(D, L, R, T, B, **kw):
            serinv_np.pobts(D, L, B, **kw)

    D, L = dense.bt_dense_to_arrays(A, b, n)
    R, T = np.zeros((n, 0, b)), np.zeros((0, 0))
    buf = np.zeros_like(D)
    selinv_permuted(D, L, R, T, buf, BT, zeros)
    Dr, Lr = dense.bt_dense_to_arrays(X, b, n)
    assert relerr(D, Dr) < RTOL and relerr(L, Lr) < RTOL
    Bm = rng.uniform(-1, 1, (n * b, 2))
    D, L = dense.bt_dense_to_arrays(A, b, n)
    buf = np.zeros_like(D)
    B = Bm.copy()
    solve_permuted(D, L, R, T, buf, B, BT, zeros)
    assert relerr(B, np.linalg.solve(A, Bm)) < RTOL


def test_permuted_workload_n8_b32_a8_vs_dense():
    b, a, n = 32, 8, 8
    A = case(b, a, n, seed=4242)
    X = np.linalg.inv(A)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    selinv_permuted(D, L, R, T, buf, OracleRT, zeros)
    Dr, Lr, Rr, Tr = dense.bta_dense_to_arrays(X, b, a, n)
    for x, y in ((D, Dr), (L, Lr), (R, Rr), (T, Tr)):
        assert relerr(x, y) < 1e-11


def test_permuted_streaming_is_rejected():
    D = np.zeros((3, 2, 2))
    with pytest.raises(NotImplementedError, match="permuted arrowhead"):
        serinv_np.pobtas(D, np.zeros((2, 2, 2)), np.zeros((3, 1, 2)), np.zeros((1, 1)),
                         np.zeros((7, 1)), buffer=np.zeros_like(D), device_streaming=True)
    with pytest.raises(NotImplementedError, match="Permuted arrowhead"):
        serinv_np.pobts(D, np.zeros((2, 2, 2)), np.zeros((6, 1)), buffer=np.zeros_like(D),
                        device_streaming=True)
