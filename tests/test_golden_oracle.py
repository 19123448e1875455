"""The oracle reproduces the committed dense-LAPACK golden vectors (CPU)."""

import numpy as np
import pytest

from oracle import serinv_np
from tests._helpers import golden_cases, relerr

CASES = golden_cases()
TOL = 1e-12


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden(name):
    c = CASES[name]
    n, b, a, r = (int(v) for v in c["shape"])
    D, L, R, T = (c[k].copy() for k in ("D", "L", "R", "T"))
    serinv_np.pobtaf(D, L, R, T)
    for got, key in zip((D, L, R, T), ("Lf_D", "Lf_L", "Lf_R", "Lf_T")):
        assert relerr(got, c[key]) < TOL, key
    assert abs(serinv_np.logdet_from_factor(D, T) - float(c["logdet"])) < TOL * max(1, abs(float(c["logdet"])))
    B = c["B"].copy()
    serinv_np.pobtas(D, L, R, T, B, trans="N")
    assert relerr(B, c["Y"]) < TOL
    serinv_np.pobtas(D, L, R, T, B, trans="C")
    assert relerr(B, c["Xs"]) < TOL
    serinv_np.pobtasi(D, L, R, T)
    for got, key in zip((D, L, R, T), ("X_D", "X_L", "X_R", "X_T")):
        assert relerr(got, c[key]) < TOL, key
