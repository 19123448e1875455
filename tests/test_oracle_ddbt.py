"""Pin the oracle's non-SPD family (ddbtasc / ddbtasci, ddbtsc / ddbtsci).

serinv's tests check, on diagonally dominant non-symmetric matrices, the
tip (or the last / first diagonal block) of the Schur-complement sweep
against the dense inverse, and the whole selected inverse after the
inversion sweep (tests/tests_algs/regular/tests_bta/test_ddbtasc.py:32-103,
test_ddbtasci.py:32-120, tests_bt/test_ddbtsc.py, test_ddbtsci.py), over
b, a in {2, 3}, n in {1..4}.  These tests do that over the same grid and
pin every block the sweep leaves: D[i] is the inverse of the Schur
complement of the leading blocks, Rl[i] and Ru[i] are that complement's
arrow blocks (from dense linear algebra on the leading rows), plus the
invert_last_block=False form and the upward BT sweeps.
"""

import numpy as np
import pytest

from oracle import dense, generators, serinv_np

RTOL = 1e-12


def relerr(x, ref):
    if ref.size == 0:
        return 0.0
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def sweep_reference(A, b, a, n, i):
    """Schur complement of A after eliminating diagonal blocks 0 .. i-1,
    restricted to block i and the tip: returns (S_ii, S_ti, S_it, S_tt)."""
    idx_i = np.arange(i * b, (i + 1) * b)
    tip = np.arange(n * b, n * b + a)
    rest = np.concatenate([np.arange(i * b, n * b), tip])  # blocks >= i and tip
    sub = A[np.ix_(rest, rest)] - A[np.ix_(rest, np.arange(i * b))] @ np.linalg.solve(
        A[:i * b, :i * b], A[np.ix_(np.arange(i * b), rest)]) if i else A[np.ix_(rest, rest)]
    m = len(rest)
    bi = np.arange(0, b)
    ti = np.arange(m - a, m)
    return (sub[np.ix_(bi, bi)], sub[np.ix_(ti, bi)], sub[np.ix_(bi, ti)], sub[np.ix_(ti, ti)])


GRID = [(b, a, n) for b in (2, 3) for a in (0, 2, 3) for n in (1, 2, 3, 4)]


@pytest.mark.parametrize("invert_last_block", [True, False])
@pytest.mark.parametrize("b,a,n", GRID)
def test_ddbtasc_blocks_vs_dense_schur(b, a, n, invert_last_block):
    A = generators.dd_bta(b, a, n, np.random.default_rng(31000 + 100 * b + 10 * a + n))
    D, Lo, Up, Rl, Ru, T = dense.bta_dense_to_arrays_full(A, b, a, n)
    Lo0, Up0 = Lo.copy(), Up.copy()
    serinv_np.ddbtasc(D, Lo, Up, Rl, Ru, T, invert_last_block=invert_last_block)
    assert np.array_equal(Lo, Lo0) and np.array_equal(Up, Up0)
    for i in range(n):
        Sii, Sti, Sit, Stt = sweep_reference(A, b, a, n, i)
        last = i == n - 1
        ref_D = np.linalg.inv(Sii) if (not last or invert_last_block) else Sii
        assert relerr(D[i], ref_D) < RTOL
        assert relerr(Rl[i], Sti) < RTOL
        assert relerr(Ru[i], Sit) < RTOL
    X = np.linalg.inv(A)
    if invert_last_block:
        assert relerr(T, X[n * b:, n * b:]) < RTOL  # test_ddbtasc.py:103
    else:  # the tip after eliminating blocks 0 .. n-2
        assert relerr(T, sweep_reference(A, b, a, n, n - 1)[3]) < RTOL


@pytest.mark.parametrize("b,a,n", GRID)
def test_ddbtasci_vs_dense_inverse(b, a, n):
    A = generators.dd_bta(b, a, n, np.random.default_rng(32000 + 100 * b + 10 * a + n))
    X = np.linalg.inv(A)
    arrs = dense.bta_dense_to_arrays_full(A, b, a, n)
    serinv_np.ddbtasc(*arrs)
    serinv_np.ddbtasci(*arrs)
    for x, y in zip(arrs, dense.bta_dense_to_arrays_full(X, b, a, n)):
        assert relerr(x, y) < RTOL


@pytest.mark.parametrize("b,a,n", [(3, 2, 4), (2, 3, 3)])
def test_ddbtasci_partial_last_block(b, a, n):
    """invert_last_block=False: the caller supplies X_{n-1,n-1}, X_{t,n-1},
    X_{n-1,t}, X_T (ddbtasci.py:207)."""
    A = generators.dd_bta(b, a, n, np.random.default_rng(33000 + n))
    X = np.linalg.inv(A)
    D, Lo, Up, Rl, Ru, T = dense.bta_dense_to_arrays_full(A, b, a, n)
    serinv_np.ddbtasc(D, Lo, Up, Rl, Ru, T)
    Xd, _, _, Xrl, Xru, Xt = dense.bta_dense_to_arrays_full(X, b, a, n)
    D[-1], Rl[-1], Ru[-1], T[:] = Xd[-1], Xrl[-1], Xru[-1], Xt
    serinv_np.ddbtasci(D, Lo, Up, Rl, Ru, T, invert_last_block=False)
    for x, y in zip((D, Lo, Up, Rl, Ru, T), dense.bta_dense_to_arrays_full(X, b, a, n)):
        assert relerr(x, y) < RTOL


@pytest.mark.parametrize("direction", ["downward", "upward"])
@pytest.mark.parametrize("b,n", [(b, n) for b in (2, 3) for n in (1, 2, 3, 4)])
def test_ddbtsc_ddbtsci_vs_dense(b, n, direction):
    A = generators.dd_bt(b, n, np.random.default_rng(34000 + 10 * b + n))
    X = np.linalg.inv(A)
    D, Lo, Up, _, _, _ = dense.bta_dense_to_arrays_full(A, b, 0, n)
    serinv_np.ddbtsc(D, Lo, Up, direction=direction)
    # the block each sweep ends on is the inverse's (test_ddbtsc.py:84-87)
    end = -1 if direction == "downward" else 0
    assert relerr(D[end], X[(n - 1) * b:, (n - 1) * b:] if end == -1 else X[:b, :b]) < RTOL
    # every diagonal block: inverse of the Schur complement of the blocks
    # already eliminated (above for downward, below for upward)
    for i in range(n):
        blk = np.arange(i * b, (i + 1) * b)
        elim = np.arange(0, i * b) if direction == "downward" else np.arange((i + 1) * b, n * b)
        S = A[np.ix_(blk, blk)]
        if elim.size:
            S = S - A[np.ix_(blk, elim)] @ np.linalg.solve(A[np.ix_(elim, elim)], A[np.ix_(elim, blk)])
        assert relerr(D[i], np.linalg.inv(S)) < RTOL
    serinv_np.ddbtsci(D, Lo, Up, direction=direction)
    for x, y in zip((D, Lo, Up), dense.bta_dense_to_arrays_full(X, b, 0, n)[:3]):
        assert relerr(x, y) < RTOL


def test_ddbt_workload_n8_b32_a8():
    b, a, n = 32, 8, 8
    A = generators.dd_bta(b, a, n, np.random.default_rng(35000))
    X = np.linalg.inv(A)
    arrs = dense.bta_dense_to_arrays_full(A, b, a, n)
    serinv_np.ddbtasc(*arrs)
    serinv_np.ddbtasci(*arrs)
    for x, y in zip(arrs, dense.bta_dense_to_arrays_full(X, b, a, n)):
        assert relerr(x, y) < 1e-11


@pytest.mark.parametrize("b,a,n", [(b, a, n) for b in (2, 3) for a in (0, 2, 3) for n in (1, 2, 3, 4)])
def test_ddbtasc_ddbtasci_quadratic_vs_dense(b, a, n):
    """A X A^T = B (test_ddbtasci.py:37-60, 108-121): after both sweeps the
    A arrays hold A^{-1} and the B arrays A^{-1} B A^{-T} on the pattern.
    B is symmetric, as the reference's tests make it: its recurrences use
    B_ii^T in places of B_ii (ddbtasci.py:540-544)."""
    rng = np.random.default_rng(36000 + 100 * b + 10 * a + n)
    A = generators.dd_bta(b, a, n, rng)
    B = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(B)
    X = np.linalg.inv(A)
    XBX = X @ B @ X.T
    arrs = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    brr = list(dense.bta_dense_to_arrays_full(B, b, a, n))
    serinv_np.ddbtasc_quadratic(*arrs, brr[0], brr[1], brr[2], brr[3], brr[4], brr[5])
    # the tip after the sweep (test_ddbtasc.py:108-121)
    if a:
        assert relerr(brr[5], XBX[n * b:, n * b:]) < RTOL
    serinv_np.ddbtasci_quadratic(*arrs, brr[0], brr[1], brr[2], brr[3], brr[4], brr[5])
    for x, y in zip(arrs, dense.bta_dense_to_arrays_full(X, b, a, n)):
        assert relerr(x, y) < RTOL
    for x, y in zip(brr, dense.bta_dense_to_arrays_full(XBX, b, a, n)):
        assert relerr(x, y) < RTOL



@pytest.mark.parametrize("direction", ["downward", "upward"])
@pytest.mark.parametrize("b,n", [(b, n) for b in (2, 3) for n in (1, 2, 3, 4, 5)])
def test_ddbtsc_ddbtsci_quadratic_vs_dense(b, n, direction):
    """BT A X A^T = B in both directions (test_ddbtsc.py / test_ddbtsci.py,
    "AXA.T=B" x {downward, upward}): the block each sweep ends on holds
    X_ee and (X B X^T)_ee; after the selected inversion every block on the
    pattern is X's and X B X^T's."""
    rng = np.random.default_rng(37000 + 10 * b + n + (500 if direction == "upward" else 0))
    A = generators.dd_bt(b, n, rng)
    B = generators.dd_bt(b, n, rng)
    generators.symmetrize(B)
    X = np.linalg.inv(A)
    XBX = X @ B @ X.T
    D, Lo, Up, _, _, _ = dense.bta_dense_to_arrays_full(A, b, 0, n)
    BD, BLo, BUp, _, _, _ = dense.bta_dense_to_arrays_full(B, b, 0, n)
    serinv_np.ddbtsc_quadratic(D, Lo, Up, BD, BLo, BUp, direction=direction)
    e = slice((n - 1) * b, n * b) if direction == "downward" else slice(0, b)
    end = -1 if direction == "downward" else 0
    assert relerr(D[end], X[e, e]) < RTOL
    assert relerr(BD[end], XBX[e, e]) < RTOL
    serinv_np.ddbtsci_quadratic(D, Lo, Up, BD, BLo, BUp, direction=direction)
    for x, y in zip((D, Lo, Up), dense.bta_dense_to_arrays_full(X, b, 0, n)[:3]):
        assert relerr(x, y) < RTOL
    for x, y in zip((BD, BLo, BUp), dense.bta_dense_to_arrays_full(XBX, b, 0, n)[:3]):
        assert relerr(x, y) < RTOL


@pytest.mark.parametrize("b,n", [(3, 4), (5, 6)])
def test_ddbtsc_quadratic_upward_is_reversed_downward(b, n):
    """The upward sweep on A equals the downward sweep on the block-reversed
    matrix J A J, block for block (lower and upper blocks trade places);
    the selected inversions agree the same way.  Holds for the A and B
    arrays at every intermediate state, so a wrong index or a swapped
    lower/upper block in either direction fails here."""
    rng = np.random.default_rng(38000 + b + n)
    A = generators.dd_bt(b, n, rng)
    B = generators.dd_bt(b, n, rng)
    generators.symmetrize(B)
    up = [x.copy() for x in dense.bta_dense_to_arrays_full(A, b, 0, n)[:3]] + \
        [x.copy() for x in dense.bta_dense_to_arrays_full(B, b, 0, n)[:3]]
    dn = [up[0][::-1].copy(), up[2][::-1].copy(), up[1][::-1].copy(),
          up[3][::-1].copy(), up[5][::-1].copy(), up[4][::-1].copy()]
    serinv_np.ddbtsc_quadratic(*up, direction="upward")
    serinv_np.ddbtsc_quadratic(*dn, direction="downward")
    for x, y in zip(up, (dn[0][::-1], dn[2][::-1], dn[1][::-1], dn[3][::-1], dn[5][::-1],
                         dn[4][::-1])):
        assert relerr(x, y) < 1e-13
    serinv_np.ddbtsci_quadratic(*up, direction="upward")
    serinv_np.ddbtsci_quadratic(*dn, direction="downward")
    for x, y in zip(up, (dn[0][::-1], dn[2][::-1], dn[1][::-1], dn[3][::-1], dn[5][::-1],
                         dn[4][::-1])):
        assert relerr(x, y) < 1e-12


@pytest.mark.parametrize("b,n", [(2, 3), (4, 5)])
def test_ddbtsc_quadratic_downward_equals_arrow_form(b, n):
    """The BT restatement (ddbtsc.py:271-311, ddbtsci.py:299-369) against the
    BTA restatement with an empty arrow (ddbtasc.py:335-475,
    ddbtasci.py:343-576), which the BT forms specialise."""
    rng = np.random.default_rng(39000 + b + n)
    A = generators.dd_bt(b, n, rng)
    B = generators.dd_bt(b, n, rng)
    generators.symmetrize(B)
    bt = [x.copy() for x in dense.bta_dense_to_arrays_full(A, b, 0, n)[:3]] + \
        [x.copy() for x in dense.bta_dense_to_arrays_full(B, b, 0, n)[:3]]
    ar = list(dense.bta_dense_to_arrays_full(A, b, 0, n)) + list(dense.bta_dense_to_arrays_full(B, b, 0, n))
    serinv_np.ddbtsc_quadratic(*bt)
    serinv_np.ddbtasc_quadratic(*ar)
    for x, y in zip(bt, ar[:3] + ar[6:9]):
        assert relerr(x, y) < 1e-13
    serinv_np.ddbtsci_quadratic(*bt)
    serinv_np.ddbtasci_quadratic(*ar)
    for x, y in zip(bt, ar[:3] + ar[6:9]):
        assert relerr(x, y) < 1e-12
