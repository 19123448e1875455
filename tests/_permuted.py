"""The reduced-system procedure of serinv's permuted tests, for any backend.

serinv checks its permuted ("middle partition") routines by completing them
with the natural routines on the 2-block reduced system formed by block 0,
block n-1 and the tip (tests/tests_algs/permuted/test_bta/
test_pobtasi_permuted.py:79-160, test_pobtas_permuted.py:58-160).  The
functions here run that procedure with routines passed in, so that the same
steps drive the oracle (numpy arrays) and the GPU routines (torch tensors).
Arrays are indexed and assigned only, so both array types work; `zeros`
makes a new array of the caller's type and device.

Also: the dense reference for the permuted factorization, the Cholesky of
the block-permuted matrix with blocks ordered 1 .. n-2, 0, n-1, tip.
"""

import numpy as np


def reduced_from(D, R, T, buffer, zeros, b, a):
    """Reduced system (2 diagonal blocks) from the permuted factor's output
    (test_pobtasi_permuted.py:111-133)."""
    rD = zeros((2, b, b))
    rL = zeros((1, b, b))
    rR = zeros((2, a, b))
    rT = zeros((a, a))
    rD[0] = D[0]
    rD[1] = D[-1]
    rL[0] = buffer[-1].T
    rR[0] = R[0]
    rR[1] = R[-1]
    rT[:, :] = T
    return rD, rL, rR, rT


def map_back(D, R, T, buffer, rD, rL, rR, rT):
    """Reduced system's blocks back into the full arrays
    (test_pobtasi_permuted.py:150-156)."""
    D[0] = rD[0]
    D[-1] = rD[1]
    buffer[-1] = rL[0].T
    R[0] = rR[0]
    R[-1] = rR[1]
    T[:, :] = rT


def selinv_permuted(D, L, R, T, buffer, rt, zeros):
    """pobtaf(buffer=) -> reduced pobtaf + pobtasi -> map back ->
    pobtasi(buffer=).  rt: object with pobtaf, pobtasi (BTA signatures)."""
    b, a = D.shape[1], R.shape[1]
    rt.pobtaf(D, L, R, T, buffer=buffer)
    rD, rL, rR, rT = reduced_from(D, R, T, buffer, zeros, b, a)
    rt.pobtaf(rD, rL, rR, rT)
    rt.pobtasi(rD, rL, rR, rT)
    map_back(D, R, T, buffer, rD, rL, rR, rT)
    rt.pobtasi(D, L, R, T, buffer=buffer)


def solve_permuted(D, L, R, T, buffer, B, rt, zeros):
    """pobtaf(buffer=) -> reduced pobtaf -> map back -> pobtas(buffer=, "N")
    -> reduced pobtas "N" and "C" -> map back -> pobtas(buffer=, "C")
    (test_pobtas_permuted.py:58-160).  Returns nothing; B holds A^{-1} B."""
    n, b, a = D.shape[0], D.shape[1], R.shape[1]
    rt.pobtaf(D, L, R, T, buffer=buffer)
    rD, rL, rR, rT = reduced_from(D, R, T, buffer, zeros, b, a)
    rt.pobtaf(rD, rL, rR, rT)
    map_back(D, R, T, buffer, rD, rL, rR, rT)
    rt.pobtas(D, L, R, T, B, buffer=buffer, trans="N")
    rB = zeros((2 * b + a, B.shape[1]))
    rB[:b] = B[:b]
    rB[b:2 * b] = B[(n - 1) * b:n * b]
    rB[2 * b:] = B[n * b:]
    rt.pobtas(rD, rL, rR, rT, rB, trans="N")
    rt.pobtas(rD, rL, rR, rT, rB, trans="C")
    B[:b] = rB[:b]
    B[(n - 1) * b:n * b] = rB[b:2 * b]
    B[n * b:] = rB[2 * b:]
    rt.pobtas(D, L, R, T, B, buffer=buffer, trans="C")


def permutation(n, b, a):
    """Row order of the block-permuted matrix: blocks 1 .. n-2, 0, n-1, tip."""
    order = list(range(1, n - 1)) + [0, n - 1]
    idx = [np.arange(i * b, (i + 1) * b) for i in order]
    idx.append(np.arange(n * b, n * b + a))
    return np.concatenate(idx)


def dense_permuted_factor(A, n, b, a):
    """Expected output of pobtaf(buffer=) from dense linear algebra.

    With P the block permutation above and m = (n-2) b, the leading m columns
    of chol(P A P^T) are the middle blocks' factor, and the trailing
    (2b + a) x (2b + a) Schur complement S is the reduced system.  Returns a
    dict of block arrays: D (middle blocks lower factor, blocks 0 and n-1
    from S), L, R, T, buffer[1:] as pobtaf(buffer=) leaves them.
    """
    p = permutation(n, b, a)
    Ap = A[np.ix_(p, p)]
    m = (n - 2) * b
    L11 = np.linalg.cholesky(Ap[:m, :m]) if m else np.zeros((0, 0))
    L21 = np.linalg.solve(L11, Ap[m:, :m].T).T if m else np.zeros((2 * b + a, 0))
    S = Ap[m:, m:] - L21 @ L21.T
    D = np.zeros((n, b, b))
    L = np.zeros((n - 1, b, b))
    R = np.zeros((n, a, b))
    buf = np.zeros((n, b, b))
    for i in range(1, n - 1):  # middle block i sits at permuted position i - 1
        c = slice((i - 1) * b, i * b)
        D[i] = L11[c, c]
        # row block i+1: a middle block at position i, or block n-1 in S
        L[i] = L11[i * b:(i + 1) * b, c] if i + 1 < n - 1 else L21[b:2 * b, c]
        buf[i] = L21[:b, c]
        R[i] = L21[2 * b:, c]
    L[0] = A[b:2 * b, :b]  # not touched by the permuted factorization
    D[0] = S[:b, :b]
    D[n - 1] = S[b:2 * b, b:2 * b]
    buf[n - 1] = S[:b, b:2 * b]
    R[0] = S[2 * b:, :b]
    R[n - 1] = S[2 * b:, b:2 * b]
    T = S[2 * b:, 2 * b:]
    return {"D": D, "L": L, "R": R, "T": T, "buffer": buf}
