"""Pin the NumPy oracle against dense linear algebra (CPU only).

The reference ships no golden vectors; its own tests compare each routine
with a dense NumPy computation on small diagonally dominant matrices
(tests/tests_algs/regular/tests_bta/test_pobtaf.py:35,80-84,
test_pobtasi.py:35,88-91, test_pobtas.py:43,81, and the BT versions
tests_bt/test_pobtf.py, test_pobtsi.py, test_pobts.py).  These tests do the
same over the reference's parameter grid (n in 1..4, b in {2,3}, a in {2,3},
r in {1,2,3}) and at the n=32, b=64, a=16 workload, with tolerances tight
enough that a dropped term, wrong sign or transposed operand fails
(test_oracle_mutations.py shows that they do).
"""

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import dense, generators, serinv_np
from serinv_amd.utils.matrices import bta_spd, bta_flops, bta_solve_flops

RTOL = 1e-12


def relerr(x, ref):
    den = np.max(np.abs(ref)) if ref.size else 1.0
    return float(np.max(np.abs(x - ref)) / den) if ref.size else 0.0


def make_case(b, a, n, seed):
    rng = np.random.default_rng(seed)
    A = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(A)
    return A


GRID = [(b, a, n) for b in (2, 3) for a in (2, 3) for n in (1, 2, 3, 4)]


def check_factor(A, b, a, n, D, L, R, T):
    Lref = np.linalg.cholesky(A)
    Dr, Lr, Rr, Tr = dense.bta_dense_to_arrays(Lref, b, a, n)
    assert relerr(D, Dr) < RTOL
    assert relerr(L, Lr) < RTOL
    assert relerr(R, Rr) < RTOL
    assert relerr(T, Tr) < RTOL
    # upper triangles are zero-filled like scipy.linalg.cholesky(lower=True)
    assert np.all(np.triu(D, 1) == 0) and np.all(np.triu(T, 1) == 0)


def check_selinv(A, b, a, n, D, L, R, T):
    X = np.linalg.inv(A)
    Dr, Lr, Rr, Tr = dense.bta_dense_to_arrays(X, b, a, n)
    assert relerr(D, Dr) < RTOL
    assert relerr(L, Lr) < RTOL
    assert relerr(R, Rr) < RTOL
    assert relerr(T, Tr) < RTOL


@pytest.mark.parametrize("b,a,n", GRID)
def test_pobtaf_vs_dense_cholesky(b, a, n):
    A = make_case(b, a, n, seed=1000 + 100 * b + 10 * a + n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    check_factor(A, b, a, n, D, L, R, T)


@pytest.mark.parametrize("b,a,n", GRID)
def test_pobtasi_vs_dense_inverse(b, a, n):
    A = make_case(b, a, n, seed=2000 + 100 * b + 10 * a + n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    serinv_np.pobtasi(D, L, R, T)
    check_selinv(A, b, a, n, D, L, R, T)


@pytest.mark.parametrize("r", (1, 2, 3))
@pytest.mark.parametrize("b,a,n", GRID)
def test_pobtas_vs_dense_solve(b, a, n, r):
    rng = np.random.default_rng(3000 + 100 * b + 10 * a + n + 7 * r)
    A = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(A)
    Bm = generators.rhs(r, b, a, n, rng)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    B = Bm.copy()
    serinv_np.pobtas(D, L, R, T, B, trans="N")
    # forward half = dense lower-triangular solve with the dense factor
    Lref = np.linalg.cholesky(A)
    assert relerr(B, sla.solve_triangular(Lref, Bm, lower=True)) < RTOL
    serinv_np.pobtas(D, L, R, T, B, trans="C")
    assert relerr(B, np.linalg.solve(A, Bm)) < RTOL


@pytest.mark.parametrize("b,n", [(b, n) for b in (2, 3) for n in (1, 2, 3, 4)])
def test_bt_routines_vs_dense(b, n):
    rng = np.random.default_rng(4000 + 10 * b + n)
    A = generators.dd_bt(b, n, rng)
    generators.symmetrize(A)
    D, L = dense.bt_dense_to_arrays(A, b, n)
    serinv_np.pobtf(D, L)
    Lref = np.linalg.cholesky(A)
    Dr, Lr = dense.bt_dense_to_arrays(Lref, b, n)
    assert relerr(D, Dr) < RTOL and relerr(L, Lr) < RTOL
    for r in (1, 2, 3):
        Bm = generators.rhs(r, b, 0, n, rng)
        B = Bm.copy()
        serinv_np.pobts(D, L, B, trans="N")
        assert relerr(B, sla.solve_triangular(Lref, Bm, lower=True)) < RTOL
        serinv_np.pobts(D, L, B, trans="C")
        assert relerr(B, np.linalg.solve(A, Bm)) < RTOL
    serinv_np.pobtsi(D, L)
    Dx, Lx = dense.bt_dense_to_arrays(np.linalg.inv(A), b, n)
    assert relerr(D, Dx) < RTOL and relerr(L, Lx) < RTOL


def test_workload_n32_b64_a16_vs_dense():
    """The CPU-baseline workload: n=32, b=64, a=16, seed 63, SDD shifts."""
    n, b, a = 32, 64, 16
    D, L, R, T, B = bta_spd(n, b, a, n_rhs=1)
    A = dense.bta_arrays_to_dense(D, L, R, T)
    assert np.allclose(A, A.T)
    Lref = np.linalg.cholesky(A)
    sign, ld = np.linalg.slogdet(A)
    Xref = np.linalg.inv(A)
    xref = np.linalg.solve(A, B)
    serinv_np.pobtaf(D, L, R, T)
    for got, ref in zip((D, L, R, T), dense.bta_dense_to_arrays(Lref, b, a, n)):
        assert relerr(got, ref) < RTOL
    assert sign > 0 and abs(serinv_np.logdet_from_factor(D, T) - ld) < 1e-12 * abs(ld)
    Bw = B.copy()
    serinv_np.pobtas(D, L, R, T, Bw, trans="N")
    serinv_np.pobtas(D, L, R, T, Bw, trans="C")
    assert relerr(Bw, xref) < RTOL
    serinv_np.pobtasi(D, L, R, T)
    for got, ref in zip((D, L, R, T), dense.bta_dense_to_arrays(Xref, b, a, n)):
        assert relerr(got, ref) < RTOL


def test_flop_formulas_match_survey_figures():
    f_af, f_asi = bta_flops(32, 64, 16)
    assert abs(f_af - 2.574e7) / 2.574e7 < 1e-3
    assert abs(f_asi - 6.729e7) / 6.729e7 < 1e-3
    f_af, f_asi = bta_flops(256, 1024, 32)
    assert abs(f_af + f_asi - 2.461e12) / 2.461e12 < 1e-3
    assert abs(2 * bta_solve_flops(32, 64, 16, 1) - 9.0e5) / 9.0e5 < 1e-2


@pytest.mark.parametrize("b,a,n", [(3, 2, 4), (2, 3, 3), (5, 4, 6)])
def test_partial_factorization_is_dense_schur_complement(b, a, n):
    """factorize_last_block=False leaves the Schur complement of the first n-1
    block rows in D[-1], R[-1], T (src/serinv/algs/pobtaf.py:161)."""
    A = make_case(b, a, n, seed=5000 + n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T, factorize_last_block=False)
    P = slice(0, (n - 1) * b)
    Q = slice((n - 1) * b, n * b + a)
    S = A[Q, Q] - A[Q, P] @ np.linalg.solve(A[P, P], A[P, Q])
    assert relerr(D[-1], S[:b, :b]) < RTOL
    assert relerr(R[-1], S[b:, :b]) < RTOL
    assert relerr(T, S[b:, b:]) < RTOL
    Lref = np.linalg.cholesky(A)
    Dr, Lr, Rr, _ = dense.bta_dense_to_arrays(Lref, b, a, n)
    assert relerr(D[:-1], Dr[:-1]) < RTOL
    assert relerr(L, Lr) < RTOL
    assert relerr(R[:-1], Rr[:-1]) < RTOL


@pytest.mark.parametrize("b,a,n", [(3, 2, 4), (2, 3, 3), (5, 4, 6)])
def test_partial_selected_inversion(b, a, n):
    """invert_last_block=False takes X_{n-1,n-1}, X_{t,n-1}, X_T as given
    (src/serinv/algs/pobtasi.py:114-141)."""
    A = make_case(b, a, n, seed=6000 + n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    Xd, Xl, Xr, Xt = dense.bta_dense_to_arrays(np.linalg.inv(A), b, a, n)
    D[-1], R[-1], T[:] = Xd[-1], Xr[-1], Xt
    serinv_np.pobtasi(D, L, R, T, invert_last_block=False)
    for got, ref in zip((D, L, R, T), (Xd, Xl, Xr, Xt)):
        assert relerr(got, ref) < RTOL


@pytest.mark.parametrize("b,a,n", [(3, 2, 4), (4, 3, 5)])
def test_partial_solves(b, a, n):
    """partial=True skips the last-block and tip solves (pobtas.py:99,129)."""
    rng = np.random.default_rng(7000 + n)
    A = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(A)
    Bm = generators.rhs(2, b, a, n, rng)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    Lref = np.linalg.cholesky(A)
    P = slice(0, (n - 1) * b)
    Q = slice((n - 1) * b, n * b + a)
    B = Bm.copy()
    serinv_np.pobtas(D, L, R, T, B, trans="N", partial=True)
    yP = sla.solve_triangular(Lref[P, P], Bm[P], lower=True)
    assert relerr(B[P], yP) < RTOL
    assert relerr(B[Q], Bm[Q] - Lref[Q, P] @ yP) < RTOL
    y = sla.solve_triangular(Lref, Bm, lower=True)
    x = np.linalg.solve(A, Bm)
    B = y.copy()
    B[Q] = x[Q]
    serinv_np.pobtas(D, L, R, T, B, trans="C", partial=True)
    assert relerr(B, x) < RTOL


def test_error_behaviour():
    D, L, R, T, B = bta_spd(3, 4, 2, n_rhs=1)
    with pytest.raises(TypeError):
        serinv_np.pobtaf(D, L, R, T, bogus=1)
    with pytest.raises(TypeError):
        serinv_np.pobtasi(D, L, R, T, bogus=1)
    with pytest.raises(ValueError):
        serinv_np.pobtas(D, L, R, T, B, trans="X")
    with pytest.raises(NotImplementedError):
        serinv_np.pobtas(D, L, R, T, B, device_streaming=True)
    D[1] = -np.eye(4)
    with pytest.raises(np.linalg.LinAlgError):
        serinv_np.pobtaf(D, L, R, T)


@pytest.mark.parametrize("b,a,n", [(2, 2, 1), (3, 3, 4), (8, 5, 7)])
def test_logdet_vs_slogdet(b, a, n):
    A = make_case(b, a, n, seed=8000 + n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    sign, ld = np.linalg.slogdet(A)
    assert sign > 0
    assert abs(serinv_np.logdet_from_factor(D, T) - ld) < 1e-12 * max(1.0, abs(ld))
