"""Seeded synthetic SPD BTA / BT matrices in serinv's block-array layout.

Construction (fill order D, L, R, T, B from one numpy Generator):
    D = U[0,1)^(n,b,b);  D = (D + D^T)/2 + gamma*(3b + a) * I
    L = U[0,1)^(n-1,b,b)
    R = U[0,1)^(n,a,b)
    T = U[0,1)^(a,a);    T = (T + T^T)/2 + gamma*(n*b + a) * I
    B = U[-1,1)^(n*b + a, r)
With gamma = 1 every row is strictly diagonally dominant (off-diagonal row
sums are below 3b + a - 1 and n*b + a - 1), so the matrix is SPD.  Smaller
gamma moves the matrix towards the PD boundary and raises its condition
number; `serinv_amd.utils.conditioning` calibrates gamma for a target kappa.

The default seed 63 is the seed of serinv's test-suite
(tests/testing_utils.py:7).
"""

import numpy as np

SEED = 63


def bta_spd(n, b, a, *, seed=SEED, gamma=1.0, n_rhs=0):
    """Return (D, L, R, T, B) as C-contiguous float64 numpy arrays.

    B is None when n_rhs == 0.  a == 0 gives the arrowless (BT) case with
    R of shape (n, 0, b) and T of shape (0, 0).
    """
    if n < 1 or b < 1 or a < 0:
        raise ValueError(f"invalid BTA shape n={n} b={b} a={a}")
    rng = np.random.default_rng(seed)
    D = rng.random((n, b, b))
    D += D.transpose(0, 2, 1).copy()
    D *= 0.5
    idx = np.arange(b)
    D[:, idx, idx] += gamma * (3 * b + a)
    L = rng.random((n - 1, b, b))
    R = rng.random((n, a, b))
    T = rng.random((a, a))
    T = (T + T.T) / 2
    T[np.arange(a), np.arange(a)] += gamma * (n * b + a)
    B = rng.uniform(-1.0, 1.0, (n * b + a, n_rhs)) if n_rhs else None
    return (np.ascontiguousarray(D), np.ascontiguousarray(L),
            np.ascontiguousarray(R), np.ascontiguousarray(T), B)


def bt_spd(n, b, *, seed=SEED, gamma=1.0, n_rhs=0):
    """Arrowless variant: (D, L, B) with the a = 0 construction."""
    D, L, _, _, B = bta_spd(n, b, 0, seed=seed, gamma=gamma, n_rhs=n_rhs)
    return D, L, B


def bta_flops(n, b, a):
    """Algorithmic FLOPs (LAPACK counting) of pobtaf and pobtasi.

    potrf m^3/3, trtri m^3/3, trsm/trmm m*k^2, syrk m^2*k, gemm 2mnk, applied
    to the reference recurrences (src/serinv/algs/pobtaf.py:115-183,
    pobtasi.py:114-172).  These counts are fixed by the recurrence, not by
    the kernels that evaluate it.
    """
    f_af = (n - 1) * (7.0 / 3.0 * b**3 + 3.0 * a * b**2 + a * a * b) + (
        b**3 / 3.0 + a * b**2 + a * a * b + a**3 / 3.0)
    f_asi = (n - 1) * (19.0 / 3.0 * b**3 + 7.0 * a * b**2 + 2.0 * a * a * b) + (
        4.0 / 3.0 * b**3 + 3.0 * a * b**2 + 2.0 * a * a * b + 2.0 / 3.0 * a**3)
    return f_af, f_asi


def bta_solve_flops(n, b, a, r):
    """FLOPs of one pobtas sweep (src/serinv/algs/pobtas.py:80-159)."""
    return r * (n * b * b + 2.0 * (n - 1) * b * b + 2.0 * n * a * b + a * a)


def bta_factor_bytes(n, b, a):
    """Bytes of one read of the stored BTA factor (D, L, R, T) in float64."""
    return 8 * (n * b * b + (n - 1) * b * b + n * a * b + a * a)
