"""Planted mistakes in copies of the oracle must fail the dense pinning checks.

Each mutant is the oracle function's own source with one plausible mistake
(a dropped term, a wrong sign, a transposed operand, a wrong index),
compiled into a private namespace.  The real oracle is never modified.
"""

import inspect
import textwrap

import numpy as np
import pytest

from oracle import dense, generators, serinv_np


def mutant(func, old, new):
    src = textwrap.dedent(inspect.getsource(func))
    assert old in src, f"mutation anchor not found: {old!r}"
    src = src.replace(old, new, 1)
    ns = dict(vars(serinv_np))
    exec(compile(src, f"<mutant {func.__name__}>", "exec"), ns)
    return ns[func.__name__]


def case(b=3, a=2, n=4, seed=11):
    rng = np.random.default_rng(seed)
    A = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(A)
    return A, b, a, n


def factor_err(fac):
    A, b, a, n = case()
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    fac(D, L, R, T)
    ref = dense.bta_dense_to_arrays(np.linalg.cholesky(A), b, a, n)
    return max(np.max(np.abs(x - y)) for x, y in zip((D, L, R, T), ref))


def selinv_err(si):
    A, b, a, n = case()
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    si(D, L, R, T)
    ref = dense.bta_dense_to_arrays(np.linalg.inv(A), b, a, n)
    return max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip((D, L, R, T), ref))


def solve_err(sv):
    A, b, a, n = case()
    rng = np.random.default_rng(5)
    Bm = rng.random((A.shape[0], 2))
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    B = Bm.copy()
    sv(D, L, R, T, B, trans="N")
    sv(D, L, R, T, B, trans="C")
    return np.max(np.abs(B - np.linalg.solve(A, Bm)))


TOL = 1e-12


def test_unmutated_oracle_passes():
    assert factor_err(serinv_np.pobtaf) < TOL
    assert selinv_err(serinv_np.pobtasi) < TOL
    assert solve_err(serinv_np.pobtas) < TOL


FACTOR_MUTANTS = [
    ("T[:, :] = T - R[i] @ R[i].conj().T", "T[:, :] = T"),  # dropped tip update
    ("R[i + 1] = R[i + 1] - R[i] @ L[i].conj().T", "R[i + 1] = R[i + 1] + R[i] @ L[i].conj().T"),
    ("D[i + 1] = D[i + 1] - L[i] @ L[i].conj().T", "D[i + 1] = D[i + 1] - L[i].conj().T @ L[i]"),
    ("L[i] = sla.solve_triangular(D[i], L[i].conj().T, lower=True).conj().T",
     "L[i] = sla.solve_triangular(D[i], L[i], lower=True)"),
]


@pytest.mark.parametrize("old,new", FACTOR_MUTANTS)
def test_factor_mutants_are_caught(old, new):
    assert factor_err(mutant(serinv_np.pobtaf, old, new)) > 1e-6


SELINV_MUTANTS = [
    ("- R[i + 1].conj().T @ Lti) @ Linv", ") @ Linv"),  # dropped arrow term
    ("R[i] = (-R[i + 1] @ Li - T @ Lti) @ Linv", "R[i] = (-R[i + 1] @ Li + T @ Lti) @ Linv"),
    ("L[i] = (-D[i + 1] @ Li", "L[i] = (-D[i] @ Li"),  # wrong index
    ("T[:, :] = LT_inv.conj().T @ LT_inv", "T[:, :] = LT_inv @ LT_inv.conj().T"),
]


@pytest.mark.parametrize("old,new", SELINV_MUTANTS)
def test_selinv_mutants_are_caught(old, new):
    assert selinv_err(mutant(serinv_np.pobtasi, old, new)) > 1e-6


SOLVE_MUTANTS = [
    ("B[tip] -= R[i] @ B[blk(i)]", "pass"),
    ("- R[i].conj().T @ B[tip]", "+ R[i].conj().T @ B[tip]"),
    ("B[blk(i + 1)] -= L[i] @ B[blk(i)]", "B[blk(i + 1)] -= L[i].conj().T @ B[blk(i)]"),
]


@pytest.mark.parametrize("old,new", SOLVE_MUTANTS)
def test_solve_mutants_are_caught(old, new):
    assert solve_err(mutant(serinv_np.pobtas, old, new)) > 1e-6


def test_logdet_mutant_is_caught():
    A, b, a, n = case()
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    _, ld = np.linalg.slogdet(A)
    bad = mutant(serinv_np.logdet_from_factor, "s += 2.0 * np.sum(np.log(np.diag(T)))", "pass")
    assert abs(bad(D, T) - ld) > 1e-6
    assert abs(serinv_np.logdet_from_factor(D, T) - ld) < 1e-10
