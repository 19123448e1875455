"""Pin the oracle's non-SPD permuted ("middle partition") forms (CPU only).

serinv's tests run the permuted Schur-complement sweep, invert the reduced
system densely, map it back, run the permuted selected inversion and
compare with the dense inverse (tests/tests_algs/permuted/test_bta/
test_ddbtasc_permuted.py, test_ddbtasci_permuted.py, test_bt/
test_ddbtsc_permuted.py, test_ddbtsci_permuted.py), over b, a in {2, 3},
n in {3, 4}.  These tests run that procedure (tests/_ddbt_permuted.py) over
the grid plus n = 2, 5 and a = 0, and pin every block the sweep leaves
against dense Schur complements in the permuted elimination order: the
middle blocks' inverses, their arrow blocks, the row-0 / column-0 fill-ins
in the buffers, and the reduced system.
"""

import numpy as np
import pytest

from oracle import dense, generators, serinv_np
from tests._ddbt_permuted import selinv_ddbt_permuted

RTOL = 1e-12


def relerr(x, ref):
    if ref.size == 0:
        return 0.0
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def blk(i, b):
    return np.arange(i * b, (i + 1) * b)


def schur(A, elim, keep):
    S = A[np.ix_(keep, keep)]
    if len(elim):
        S = S - A[np.ix_(keep, elim)] @ np.linalg.solve(A[np.ix_(elim, elim)], A[np.ix_(elim, keep)])
    return S


class OracleRT:
    @staticmethod
    def sc(arrs, Lbuf, Ubuf):
        serinv_np.ddbtasc_permuted(*arrs, Lbuf, Ubuf)

    @staticmethod
    def sci(arrs, Lbuf, Ubuf):
        serinv_np.ddbtasci_permuted(*arrs, Lbuf, Ubuf)

    @staticmethod
    def rsc(r):
        serinv_np.ddbtasc(*r)

    @staticmethod
    def rsci(r):
        serinv_np.ddbtasci(*r)


GRID = [(b, a, n) for b in (2, 3) for a in (0, 2, 3) for n in (2, 3, 4, 5)]


@pytest.mark.parametrize("b,a,n", GRID)
def test_ddbtasc_permuted_blocks_vs_dense_schur(b, a, n):
    A = generators.dd_bta(b, a, n, np.random.default_rng(52000 + 100 * b + 10 * a + n))
    D, Lo, Up, Rl, Ru, T = dense.bta_dense_to_arrays_full(A, b, a, n)
    Lbuf, Ubuf = np.zeros_like(Lo), np.zeros_like(Lo)
    Lo0, Up0 = Lo.copy(), Up.copy()
    serinv_np.ddbtasc_permuted(D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf)
    assert np.array_equal(Lo, Lo0) and np.array_equal(Up, Up0)
    tip = np.arange(n * b, n * b + a)
    for i in range(1, n - 1):  # middle block i after eliminating 1 .. i-1
        elim = np.arange(b, i * b)
        S = schur(A, elim, np.concatenate([blk(i, b), tip]))
        assert relerr(D[i], np.linalg.inv(S[:b, :b])) < RTOL
        assert relerr(Rl[i], S[b:, :b]) < RTOL
        assert relerr(Ru[i], S[:b, b:]) < RTOL
    for k in range(n - 1):  # fill-ins after eliminating 1 .. k
        elim = np.arange(b, (k + 1) * b)
        S = schur(A, elim, np.concatenate([blk(0, b), blk(k + 1, b)]))
        assert relerr(Lbuf[k], S[:b, b:]) < RTOL
        assert relerr(Ubuf[k], S[b:, :b]) < RTOL
    # the reduced system: blocks 0, n-1 and the tip after all middle blocks
    keep = np.concatenate([blk(0, b), blk(n - 1, b), tip])
    S = schur(A, np.arange(b, (n - 1) * b), keep)
    r0, r1, rt = slice(0, b), slice(b, 2 * b), slice(2 * b, 2 * b + a)
    for x, y in ((D[0], S[r0, r0]), (D[n - 1], S[r1, r1]), (T, S[rt, rt]), (Rl[0], S[rt, r0]),
                 (Rl[n - 1], S[rt, r1]), (Ru[0], S[r0, rt]), (Ru[n - 1], S[r1, rt])):
        assert relerr(x, y) < RTOL


@pytest.mark.parametrize("b,a,n", GRID)
def test_ddbtasci_permuted_vs_dense_inverse(b, a, n):
    A = generators.dd_bta(b, a, n, np.random.default_rng(53000 + 100 * b + 10 * a + n))
    X = np.linalg.inv(A)
    arrs = dense.bta_dense_to_arrays_full(A, b, a, n)
    Lbuf, Ubuf = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    selinv_ddbt_permuted(arrs, Lbuf, Ubuf, OracleRT, np.zeros)
    for x, y in zip(arrs, dense.bta_dense_to_arrays_full(X, b, a, n)):
        assert relerr(x, y) < RTOL
    for k in range(n - 1):  # buffers: X_{0,k+1}, X_{k+1,0}
        assert relerr(Lbuf[k], X[np.ix_(blk(0, b), blk(k + 1, b))]) < RTOL
        assert relerr(Ubuf[k], X[np.ix_(blk(k + 1, b), blk(0, b))]) < RTOL


def test_ddbt_permuted_workload_n8_b32_a8():
    b, a, n = 32, 8, 8
    A = generators.dd_bta(b, a, n, np.random.default_rng(54000))
    arrs = dense.bta_dense_to_arrays_full(A, b, a, n)
    Lbuf, Ubuf = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    selinv_ddbt_permuted(arrs, Lbuf, Ubuf, OracleRT, np.zeros)
    for x, y in zip(arrs, dense.bta_dense_to_arrays_full(np.linalg.inv(A), b, a, n)):
        assert relerr(x, y) < 1e-11


# --------------------------------------------------------------- A X A^T = B
def transformed(A, B, elim, keep):
    """B carried through the elimination of `elim`: M B M^T on `keep`, with
    M = [-A_KE A_EE^{-1}, I] (what the quadratic sweep's updates B~_rc +=
    F_r S B~_ii S^T F_c^T - B~_ri S^T F_c^T - F_r S B~_ic build up)."""
    if not len(elim):
        return B[np.ix_(keep, keep)]
    M = np.hstack([-A[np.ix_(keep, elim)] @ np.linalg.inv(A[np.ix_(elim, elim)]), np.eye(len(keep))])
    idx = np.concatenate([elim, keep])
    return M @ B[np.ix_(idx, idx)] @ M.T


class OracleQRT:
    sc = staticmethod(lambda arrs, L, U, brr, BL, BU:
                      serinv_np.ddbtasc_quadratic_permuted(*arrs, L, U, *brr, BL, BU))
    sci = staticmethod(lambda arrs, L, U, brr, BL, BU:
                       serinv_np.ddbtasci_quadratic_permuted(*arrs, L, U, *brr, BL, BU))
    rsc = staticmethod(lambda r, rb: serinv_np.ddbtasc_quadratic(*r, *rb))
    rsci = staticmethod(lambda r, rb: serinv_np.ddbtasci_quadratic(*r, *rb))


def quad_case(b, a, n, seed):
    rng = np.random.default_rng(seed)
    A = generators.dd_bta(b, a, n, rng)
    B = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(B)
    return A, B


@pytest.mark.parametrize("b,a,n", GRID)
def test_ddbtasc_quadratic_permuted_blocks_vs_dense(b, a, n):
    A, B = quad_case(b, a, n, 57000 + 100 * b + 10 * a + n)
    arrs = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    brr = list(dense.bta_dense_to_arrays_full(B, b, a, n))
    L, U = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    BL, BU = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    serinv_np.ddbtasc_quadratic_permuted(*arrs, L, U, *brr, BL, BU)
    BD, BLo, BUp, BRl, BRu, BT = brr
    Bref = dense.bta_dense_to_arrays_full(B, b, a, n)
    assert np.array_equal(BLo, Bref[1]) and np.array_equal(BUp[1:], Bref[2][1:])
    tip = np.arange(n * b, n * b + a)
    for i in range(1, n - 1):  # middle block i after eliminating 1 .. i-1
        elim = np.arange(b, i * b)
        keep = np.concatenate([blk(i, b), tip])
        Bt = transformed(A, B, elim, keep)
        Sinv = np.linalg.inv(schur(A, elim, blk(i, b)))
        assert relerr(BD[i], Sinv @ Bt[:b, :b] @ Sinv.T) < RTOL
        assert relerr(BRl[i], Bt[b:, :b]) < RTOL
        assert relerr(BRu[i], Bt[:b, b:]) < RTOL
    for k in range(n - 1):  # B's fill-ins after eliminating 1 .. k
        Bt = transformed(A, B, np.arange(b, (k + 1) * b), np.concatenate([blk(0, b), blk(k + 1, b)]))
        assert relerr(BL[k], Bt[:b, b:]) < RTOL
        assert relerr(BU[k], Bt[b:, :b]) < RTOL
    keep = np.concatenate([blk(0, b), blk(n - 1, b), tip])
    Bt = transformed(A, B, np.arange(b, (n - 1) * b), keep)
    r0, r1, rt = slice(0, b), slice(b, 2 * b), slice(2 * b, 2 * b + a)
    for x, y in ((BD[0], Bt[r0, r0]), (BD[n - 1], Bt[r1, r1]), (BT, Bt[rt, rt]), (BRl[0], Bt[rt, r0]),
                 (BRl[n - 1], Bt[rt, r1]), (BRu[0], Bt[r0, rt]), (BRu[n - 1], Bt[r1, rt])):
        assert relerr(x, y) < RTOL
    # A's side is the plain permuted sweep
    arrs2 = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    L2, U2 = np.zeros_like(L), np.zeros_like(U)
    serinv_np.ddbtasc_permuted(*arrs2, L2, U2)
    for x, y in zip(arrs + [L, U], arrs2 + [L2, U2]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("b,a,n", GRID)
def test_ddbtasci_quadratic_permuted_vs_dense(b, a, n):
    A, B = quad_case(b, a, n, 58000 + 100 * b + 10 * a + n)
    X = np.linalg.inv(A)
    XBX = X @ B @ X.T
    arrs = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    brr = list(dense.bta_dense_to_arrays_full(B, b, a, n))
    L, U = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    BL, BU = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    selinv_ddbt_permuted(arrs, L, U, OracleQRT, np.zeros, brr, (BL, BU))
    for x, y in zip(arrs, dense.bta_dense_to_arrays_full(X, b, a, n)):
        assert relerr(x, y) < RTOL
    for x, y in zip(brr, dense.bta_dense_to_arrays_full(XBX, b, a, n)):
        assert relerr(x, y) < RTOL
    for k in range(n - 1):
        assert relerr(BL[k], XBX[np.ix_(blk(0, b), blk(k + 1, b))]) < RTOL
        assert relerr(BU[k], XBX[np.ix_(blk(k + 1, b), blk(0, b))]) < RTOL
