"""Write the golden BTA fixtures in tests/golden/bta_golden.npz.

Each case holds the input blocks (D, L, R, T), right-hand sides B and the
expected outputs computed by DENSE LAPACK on the assembled matrix (never by
the block algorithm under test):
    Lf_*  : blocks of numpy.linalg.cholesky(A)          (pobtaf output)
    X_*   : blocks of numpy.linalg.inv(A) on the pattern (pobtasi output)
    Y     : scipy.linalg.solve_triangular(chol(A), B)    (pobtas "N" output)
    Xs    : numpy.linalg.solve(A, B)                     (pobtas "N" then "C")
    logdet: numpy.linalg.slogdet(A)[1]
Cases: the reference test-suite generator at its own grid corners
(tests/testing_utils.py:346-428 restated in oracle/generators.py) and the
seeded SPD construction of serinv_amd.utils.matrices at tile-boundary shapes.

Run from the repository root:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np
import scipy.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import dense, generators  # noqa: E402
from serinv_amd.utils.matrices import bta_spd  # noqa: E402

CASES = [
    # name, kind, n, b, a, r, seed
    ("dd_n4_b3_a2", "dd", 4, 3, 2, 3, 101),
    ("dd_n1_b2_a3", "dd", 1, 2, 3, 1, 102),
    ("dd_n3_b2_a0", "dd", 3, 2, 0, 2, 103),
    ("spd_n3_b65_a5", "spd", 3, 65, 5, 2, 104),
    ("spd_n2_b70_a17", "spd", 2, 70, 17, 1, 105),
    ("spd_n5_b20_a0", "spd", 5, 20, 0, 4, 106),
]


def make(kind, n, b, a, r, seed):
    if kind == "dd":
        rng = np.random.default_rng(seed)
        A = generators.dd_bta(b, a, n, rng)
        generators.symmetrize(A)
        B = generators.rhs(r, b, a, n, rng)
    else:
        D, L, R, T, B = bta_spd(n, b, a, seed=seed, n_rhs=r)
        A = dense.bta_arrays_to_dense(D, L, R, T)
    return A, B


def main():
    out = {}
    for name, kind, n, b, a, r, seed in CASES:
        A, B = make(kind, n, b, a, r, seed)
        D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
        C = np.linalg.cholesky(A)
        X = np.linalg.inv(A)
        sign, ld = np.linalg.slogdet(A)
        assert sign > 0
        fields = {
            "shape": np.array([n, b, a, r]),
            "D": D, "L": L, "R": R, "T": T, "B": B,
            "Y": sla.solve_triangular(C, B, lower=True),
            "Xs": np.linalg.solve(A, B),
            "logdet": np.array(ld),
        }
        for tag, M in (("Lf", C), ("X", X)):
            Db, Lb, Rb, Tb = dense.bta_dense_to_arrays(M, b, a, n)
            fields.update({f"{tag}_D": Db, f"{tag}_L": Lb, f"{tag}_R": Rb, f"{tag}_T": Tb})
        for k, v in fields.items():
            out[f"{name}/{k}"] = v
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bta_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
