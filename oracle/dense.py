"""Dense <-> block-array conversions for BTA and BT matrices.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Semantics follow serinv's test helpers: bta_dense_to_arrays
(tests/testing_utils.py:17-122), bt_dense_to_arrays (:125-191) and the BTA
arrays->dense converter (:194-271).  Only the lower pattern is kept as block
arrays; assembly mirrors it to the upper triangle (A is symmetric).
"""

import numpy as np


def bta_arrays_to_dense(D, L, R, T, symmetric=True):
    """Assemble the dense N x N matrix, N = n*b + a, from BTA block arrays.

    With symmetric=True the upper blocks are the transposes of the lower
    blocks (the SPD case); otherwise the upper pattern is left at zero,
    which is what a lower Cholesky factor looks like.
    """
    n, b = D.shape[0], D.shape[1]
    a = R.shape[1] if R is not None else 0
    N = n * b + a
    A = np.zeros((N, N), dtype=D.dtype)
    for i in range(n):
        A[i * b:(i + 1) * b, i * b:(i + 1) * b] = D[i]
        if i + 1 < n:
            A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = L[i]
            if symmetric:
                A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = L[i].conj().T
        if a:
            A[n * b:, i * b:(i + 1) * b] = R[i]
            if symmetric:
                A[i * b:(i + 1) * b, n * b:] = R[i].conj().T
    if a:
        A[n * b:, n * b:] = T
    return A


def bt_arrays_to_dense(D, L, symmetric=True):
    return bta_arrays_to_dense(D, L, None, None, symmetric=symmetric)


def bta_dense_to_arrays(A, b, a, n):
    """Extract (D, L, R, T) of the lower BTA pattern from a dense matrix."""
    D = np.zeros((n, b, b), dtype=A.dtype)
    L = np.zeros((max(n - 1, 0), b, b), dtype=A.dtype)
    R = np.zeros((n, a, b), dtype=A.dtype)
    T = np.zeros((a, a), dtype=A.dtype)
    for i in range(n):
        D[i] = A[i * b:(i + 1) * b, i * b:(i + 1) * b]
        if i + 1 < n:
            L[i] = A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b]
        if a:
            R[i] = A[n * b:n * b + a, i * b:(i + 1) * b]
    if a:
        T[:] = A[n * b:, n * b:]
    return D, L, R, T


def bt_dense_to_arrays(A, b, n):
    D, L, _, _ = bta_dense_to_arrays(A, b, 0, n)
    return D, L


def bta_dense_to_arrays_full(A, b, a, n):
    """(D, Lo, Up, Rl, Ru, T) of a general (non-symmetric) BTA matrix:
    Lo[i] = A_{i+1,i}, Up[i] = A_{i,i+1}, Rl[i] = A_{t,i}, Ru[i] = A_{i,t}
    (the arrays serinv's tests/testing_utils.py:17-122 returns for the ddbt
    routines)."""
    D = np.zeros((n, b, b), dtype=A.dtype)
    Lo = np.zeros((max(n - 1, 0), b, b), dtype=A.dtype)
    Up = np.zeros((max(n - 1, 0), b, b), dtype=A.dtype)
    Rl = np.zeros((n, a, b), dtype=A.dtype)
    Ru = np.zeros((n, b, a), dtype=A.dtype)
    T = np.zeros((a, a), dtype=A.dtype)
    for i in range(n):
        blk = slice(i * b, (i + 1) * b)
        D[i] = A[blk, blk]
        if i + 1 < n:
            nxt = slice((i + 1) * b, (i + 2) * b)
            Lo[i] = A[nxt, blk]
            Up[i] = A[blk, nxt]
        if a:
            Rl[i] = A[n * b:, blk]
            Ru[i] = A[blk, n * b:]
    if a:
        T[:] = A[n * b:, n * b:]
    return D, Lo, Up, Rl, Ru, T
