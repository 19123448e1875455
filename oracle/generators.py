"""Random diagonally dominant BT/BTA test matrices, as serinv's test-suite builds them.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates dd_bta (tests/testing_utils.py:346-428), dd_bt (:431-496), rhs
(:499-540) and symmetrize (:332-343) with an explicit numpy Generator in
place of the reference's import-time global seed (:7-9), so every test case
is reproducible on its own.  Fill order is the reference's:
lower arrow, upper arrow, tip, then per block row diagonal / lower / upper,
then A[i, i] = 1 + sum(A[i, :]) and symmetrization (A + A^H) / 2.
Only float64 is generated: the reference's complex fixture passes the string
"complex128", which never equals np.complex128 in dd_bta (:386), so its
"complex" matrices are real anyway.
"""

import numpy as np


def dd_bta(b, a, n, rng):
    N = b * n + a
    A = np.zeros((N, N))
    if a:
        A[-a:, :-a] = rng.random((a, n * b))  # :389-391
        A[:-a, -a:] = rng.random((n * b, a))  # :393-395
        A[-a:, -a:] = rng.random((a, a))  # :398-400
    for i in range(n):  # :403-422
        A[i * b:(i + 1) * b, i * b:(i + 1) * b] = rng.random((b, b)) + np.eye(b)
        if i > 0:
            A[i * b:(i + 1) * b, (i - 1) * b:i * b] = rng.random((b, b))
        if i < n - 1:
            A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = rng.random((b, b))
    for i in range(N):  # :425-426
        A[i, i] = 1 + np.sum(A[i, :])
    return A


def dd_bt(b, n, rng):
    return dd_bta(b, 0, n, rng)


def symmetrize(A):
    A[:] = (A + A.conj().T) / 2  # :343


def rhs(n_rhs, b, a, n, rng):
    return rng.random((b * n + a, n_rhs))  # :536-538
