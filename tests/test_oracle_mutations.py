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


# Permuted forms: the mutant replaces one routine in the reduced-system
# procedure, checked against the dense inverse / solve, and the mutated
# factorization also against the Cholesky of the block-permuted matrix.
from tests._permuted import dense_permuted_factor, selinv_permuted, solve_permuted  # noqa: E402


def permuted_rt(f=None, si=None, sv=None):
    class RT:
        @staticmethod
        def pobtaf(D, L, R, T, buffer=None):
            if buffer is None:
                serinv_np.pobtaf(D, L, R, T)
            else:
                (f or serinv_np.pobtaf_permuted)(D, L, R, T, buffer)

        @staticmethod
        def pobtasi(D, L, R, T, buffer=None):
            if buffer is None:
                serinv_np.pobtasi(D, L, R, T)
            else:
                (si or serinv_np.pobtasi_permuted)(D, L, R, T, buffer)

        @staticmethod
        def pobtas(D, L, R, T, B, buffer=None, trans="N"):
            if buffer is None:
                serinv_np.pobtas(D, L, R, T, B, trans=trans)
            else:
                (sv or serinv_np.pobtas_permuted)(D, L, R, T, B, buffer, trans)
    return RT


def permuted_errs(rt, n=5):
    A, b, a, n = case(n=n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    rt.pobtaf(D, L, R, T, buffer=buf)
    ref = dense_permuted_factor(A, n, b, a)
    fac = max(np.max(np.abs(x - ref[k])) for k, x in zip("DLRT", (D, L, R, T)))
    fac = max(fac, np.max(np.abs(buf[1:] - ref["buffer"][1:])))
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    selinv_permuted(D, L, R, T, buf, rt, np.zeros)
    ref = dense.bta_dense_to_arrays(np.linalg.inv(A), b, a, n)
    si = max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip((D, L, R, T), ref))
    Bm = np.random.default_rng(5).random((A.shape[0], 2))
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    buf = np.zeros_like(D)
    B = Bm.copy()
    solve_permuted(D, L, R, T, buf, B, rt, np.zeros)
    sv = np.max(np.abs(B - np.linalg.solve(A, Bm)))
    return fac, si, sv


def test_unmutated_permuted_oracle_passes():
    assert max(permuted_errs(permuted_rt())) < TOL


PERMUTED_FACTOR_MUTANTS = [
    ("D[0] = D[0] - buffer[i] @ buffer[i].conj().T", "D[0] = D[0] + buffer[i] @ buffer[i].conj().T"),
    ("buffer[i + 1] = -buffer[i] @ L[i].conj().T", "buffer[i + 1] = buffer[i] @ L[i].conj().T"),
    ("R[0] = R[0] - R[i] @ buffer[i].conj().T", "R[0] = R[0] - R[i] @ L[i].conj().T"),
    ("buffer[1] = L[0].conj().T", "buffer[1] = L[1].conj().T"),  # wrong index
]


@pytest.mark.parametrize("old,new", PERMUTED_FACTOR_MUTANTS)
def test_permuted_factor_mutants_are_caught(old, new):
    bad = mutant(serinv_np.pobtaf_permuted, old, new)
    fac, si, sv = permuted_errs(permuted_rt(f=bad))
    assert fac > 1e-6 and si > 1e-6 and sv > 1e-6


PERMUTED_SELINV_MUTANTS = [
    (" - R[0].conj().T @ Ri) @ Linv", ") @ Linv"),  # dropped arrow term
    ("-buffer[i + 1].conj().T @ Bi", "-buffer[i + 1] @ Bi"),  # transposed operand
    ("- D[0] @ Bi", "- D[i] @ Bi"),  # wrong index
    ("L[0] = buffer[1].conj().T", "pass"),
]


@pytest.mark.parametrize("old,new", PERMUTED_SELINV_MUTANTS)
def test_permuted_selinv_mutants_are_caught(old, new):
    bad = mutant(serinv_np.pobtasi_permuted, old, new)
    assert permuted_errs(permuted_rt(si=bad))[1] > 1e-6


PERMUTED_SOLVE_MUTANTS = [
    ("- buffer[i].conj().T @ B[blk(0)]", "+ buffer[i].conj().T @ B[blk(0)]"),
    ("B[blk(0)] -= buffer[i] @ B[blk(i)]", "pass"),
    ("B[tip] -= R[i] @ B[blk(i)]", "B[tip] -= R[i - 1] @ B[blk(i)]"),
]


@pytest.mark.parametrize("old,new", PERMUTED_SOLVE_MUTANTS)
def test_permuted_solve_mutants_are_caught(old, new):
    bad = mutant(serinv_np.pobtas_permuted, old, new)
    assert permuted_errs(permuted_rt(sv=bad))[2] > 1e-6


# Non-SPD family: mutants of the sweep and of the inversion must change the
# selected inverse against the dense inverse of a non-symmetric matrix.
def ddbt_err(sc=None, sci=None):
    b, a, n = 3, 2, 4
    A = generators.dd_bta(b, a, n, np.random.default_rng(77))
    arrs = dense.bta_dense_to_arrays_full(A, b, a, n)
    (sc or serinv_np.ddbtasc)(*arrs)
    (sci or serinv_np.ddbtasci)(*arrs)
    ref = dense.bta_dense_to_arrays_full(np.linalg.inv(A), b, a, n)
    return max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(arrs, ref))


def test_unmutated_ddbt_oracle_passes():
    assert ddbt_err() < TOL


DDBT_SC_MUTANTS = [
    ("Ru[i] = Ru[i] - Lo[i - 1] @ D[i - 1] @ Ru[i - 1]", "Ru[i] = Ru[i] - Up[i - 1] @ D[i - 1] @ Ru[i - 1]"),
    ("T[:, :] = T - Rl[i - 1] @ D[i - 1] @ Ru[i - 1]", "T[:, :] = T"),
    ("D[i] = D[i] - Lo[i - 1] @ D[i - 1] @ Up[i - 1]", "D[i] = D[i] - Up[i - 1] @ D[i - 1] @ Lo[i - 1]"),
]


@pytest.mark.parametrize("old,new", DDBT_SC_MUTANTS)
def test_ddbtasc_mutants_are_caught(old, new):
    assert ddbt_err(sc=mutant(serinv_np.ddbtasc, old, new)) > 1e-6


DDBT_SCI_MUTANTS = [
    ("C1 = D[i + 1] @ Lo[i] + Ru[i + 1] @ Rl[i]", "C1 = D[i + 1] @ Lo[i]"),
    ("Rl[i] = -C2 @ D[i]", "Rl[i] = C2 @ D[i]"),
    ("D[i] = D[i] + D[i] @ (B1 @ D1 + B2 @ D2) @ D[i]", "D[i] = D[i] + D[i] @ (B1 @ D1) @ D[i]"),
    ("B2 = Up[i] @ Ru[i + 1] + Ru[i] @ T", "B2 = Up[i] @ Ru[i + 1] + Ru[i] @ T.T"),
]


@pytest.mark.parametrize("old,new", DDBT_SCI_MUTANTS)
def test_ddbtasci_mutants_are_caught(old, new):
    assert ddbt_err(sci=mutant(serinv_np.ddbtasci, old, new)) > 1e-6


# BT A X A^T = B, both directions: mutants must change X or X B X^T.
def ddbt_quad_err(direction, sc=None, sci=None):
    b, n = 3, 4
    rng = np.random.default_rng(78)
    A = generators.dd_bt(b, n, rng)
    B = generators.dd_bt(b, n, rng)
    generators.symmetrize(B)
    arrs = list(dense.bta_dense_to_arrays_full(A, b, 0, n)[:3]) + \
        list(dense.bta_dense_to_arrays_full(B, b, 0, n)[:3])
    (sc or serinv_np.ddbtsc_quadratic)(*arrs, direction=direction)
    (sci or serinv_np.ddbtsci_quadratic)(*arrs, direction=direction)
    X = np.linalg.inv(A)
    ref = list(dense.bta_dense_to_arrays_full(X, b, 0, n)[:3]) + \
        list(dense.bta_dense_to_arrays_full(X @ B @ X.T, b, 0, n)[:3])
    return max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(arrs, ref))


@pytest.mark.parametrize("direction", ["downward", "upward"])
def test_unmutated_ddbt_quadratic_oracle_passes(direction):
    assert ddbt_quad_err(direction) < TOL


DDBT_QSC_MUTANTS = [
    ("upward", "BD[i - 1] + Up[i - 1] @ BD[i] @ Up[i - 1].T - BUp[i - 1] @ UD.T - UD @ BLo[i - 1]",
     "BD[i - 1] + Up[i - 1] @ BD[i] @ Up[i - 1].T - BLo[i - 1] @ UD.T - UD @ BUp[i - 1]"),
    ("upward", "D[i - 1] = D[i - 1] - UD @ Lo[i - 1]", "D[i - 1] = D[i - 1] - UD @ Up[i - 1]"),
    ("upward", "BD[last] = D[last] @ BD[last] @ D[last].T", "BD[last] = D[last] @ BD[last]"),
    ("downward", "- BLo[i] @ LD.T - LD @ BUp[i]", "- BLo[i] @ LD.T"),
]


@pytest.mark.parametrize("direction,old,new", DDBT_QSC_MUTANTS)
def test_ddbtsc_quadratic_mutants_are_caught(direction, old, new):
    bad = mutant(serinv_np.ddbtsc_quadratic, old, new)
    assert ddbt_quad_err(direction, sc=bad) > 1e-6


_QSCI_BOTH = [
    ("t4 = AL[k].T @ D[f].T", "t4 = AL[k] @ D[f].T"),
    ("+ t1 @ t3 @ BD[g] +", "+"),
    ("BL[k] = -BD[f] @ t2 - t3 @ BD[g] + D[f] @ Bfg @ D[g].T",
     "BL[k] = -BD[f] @ t2 - t3 @ BD[g] + D[f] @ Bgf @ D[g].T"),
]
DDBT_QSCI_MUTANTS = [(d, o, nw) for d in ("downward", "upward") for o, nw in _QSCI_BOTH] + [
    ("upward", "AU, AL, BU, BL = Lo, Up, BLo, BUp", "AU, AL, BU, BL = Lo, Up, BUp, BLo"),
]


@pytest.mark.parametrize("direction,old,new", DDBT_QSCI_MUTANTS)
def test_ddbtsci_quadratic_mutants_are_caught(direction, old, new):
    bad = mutant(serinv_np.ddbtsci_quadratic, old, new)
    assert ddbt_quad_err(direction, sci=bad) > 1e-6


# Non-SPD permuted forms (A X = I and A X A^T = B): mutants must change the
# completed inverse (or A^{-1} B A^{-T}) against dense linear algebra.
def ddbt_perm_err(sc=None, sci=None, quadratic=False):
    from tests._ddbt_permuted import selinv_ddbt_permuted

    b, a, n = 3, 2, 5
    rng = np.random.default_rng(79)
    A = generators.dd_bta(b, a, n, rng)
    B = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(B)
    arrs = list(dense.bta_dense_to_arrays_full(A, b, a, n))
    L, U = np.zeros_like(arrs[1]), np.zeros_like(arrs[1])
    X = np.linalg.inv(A)
    if not quadratic:
        class RT:
            sc = staticmethod(lambda ar, l, u: (sc or serinv_np.ddbtasc_permuted)(*ar, l, u))
            sci = staticmethod(lambda ar, l, u: (sci or serinv_np.ddbtasci_permuted)(*ar, l, u))
            rsc = staticmethod(lambda r: serinv_np.ddbtasc(*r))
            rsci = staticmethod(lambda r: serinv_np.ddbtasci(*r))
        selinv_ddbt_permuted(arrs, L, U, RT, np.zeros)
        got, ref = arrs, dense.bta_dense_to_arrays_full(X, b, a, n)
    else:
        brr = list(dense.bta_dense_to_arrays_full(B, b, a, n))
        BL, BU = np.zeros_like(L), np.zeros_like(U)

        class RT:
            sc = staticmethod(lambda ar, l, u, br, bl, bu:
                              (sc or serinv_np.ddbtasc_quadratic_permuted)(*ar, l, u, *br, bl, bu))
            sci = staticmethod(lambda ar, l, u, br, bl, bu:
                               (sci or serinv_np.ddbtasci_quadratic_permuted)(*ar, l, u, *br, bl, bu))
            rsc = staticmethod(lambda r, rb: serinv_np.ddbtasc_quadratic(*r, *rb))
            rsci = staticmethod(lambda r, rb: serinv_np.ddbtasci_quadratic(*r, *rb))
        selinv_ddbt_permuted(arrs, L, U, RT, np.zeros, brr, (BL, BU))
        got = arrs + brr
        ref = list(dense.bta_dense_to_arrays_full(X, b, a, n)) + \
            list(dense.bta_dense_to_arrays_full(X @ B @ X.T, b, a, n))
    return max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(got, ref))


@pytest.mark.parametrize("quadratic", [False, True])
def test_unmutated_ddbt_permuted_oracle_passes(quadratic):
    assert ddbt_perm_err(quadratic=quadratic) < TOL


DDBT_PERM_MUTANTS = [
    ("sc", "Ubuf[i] = -Lo[i] @ S @ Ubuf[i - 1]", "Ubuf[i] = -Up[i] @ S @ Ubuf[i - 1]"),
    ("sc", "Ru[0] = Ru[0] - Lbuf[i - 1] @ S @ Ru[i]", "Ru[0] = Ru[0]"),
    ("sci", "C2 = Lbuf[i] @ Lo[i] + D[0] @ Lbuf[i - 1] + Ru[0] @ Rl[i]",
     "C2 = Lbuf[i] @ Lo[i] + D[0] @ Lbuf[i - 1]"),
    ("sci", "B3 = Up[i] @ Ru[i + 1] + Ru[i] @ T + Ubuf[i - 1] @ Ru[0]",
     "B3 = Up[i] @ Ru[i + 1] + Ru[i] @ T"),
    ("sci", "        Lo[0] = Ubuf[0]  # X_{1,0}", "        Lo[0] = Lbuf[0]  # X_{1,0}"),
]


@pytest.mark.parametrize("which,old,new", DDBT_PERM_MUTANTS)
def test_ddbt_permuted_mutants_are_caught(which, old, new):
    if which == "sc":
        err = ddbt_perm_err(sc=mutant(serinv_np.ddbtasc_permuted, old, new))
    else:
        err = ddbt_perm_err(sci=mutant(serinv_np.ddbtasci_permuted, old, new))
    assert err > 1e-6


DDBT_QPERM_MUTANTS = [
    ("sc", "BUbuf[i] = BUbuf[i] + upd(\"n\", \"0\")", "BUbuf[i] = BUbuf[i] + upd(\"0\", \"n\")"),
    ("sc", "return F[r] @ BD[i] @ F[c].T - G[r] @ S.T @ F[c].T - F[r] @ S @ H[c]",
     "return F[r] @ BD[i] @ F[c].T - G[r] @ S.T @ F[c].T + F[r] @ S @ H[c]"),
    ("sc", "BD[i] = S @ BD[i] @ S.T  # :581-583", "BD[i] = S @ BD[i]  # :581-583"),
    ("sci", "Gj = [sum(bR[k] @ X[j][k].T for k in J) for j in J]",
     "Gj = [sum(bR[k] @ X[j][k].T for k in range(2)) for j in J]"),
    ("sci", "newC = [-DK[j] @ S.T - Cj[j] @ SB + Hj[j] @ S.T for j in J]",
     "newC = [-DK[j] @ S.T - Cj[j] @ SB + Hj[j] @ S for j in J]"),
    ("sci", "                 - S @ sum(Bj[j] @ bC[j] for j in J) @ S.T",
     "                 - S @ sum(Bj[j] @ bC[j] for j in range(2)) @ S.T"),
    ("sci", "        BLo[0] = BUbuf[0]", "        BLo[0] = BLbuf[0]"),
]


@pytest.mark.parametrize("which,old,new", DDBT_QPERM_MUTANTS)
def test_ddbt_quadratic_permuted_mutants_are_caught(which, old, new):
    if which == "sc":
        err = ddbt_perm_err(sc=mutant(serinv_np.ddbtasc_quadratic_permuted, old, new), quadratic=True)
    else:
        err = ddbt_perm_err(sci=mutant(serinv_np.ddbtasci_quadratic_permuted, old, new),
                            quadratic=True)
    assert err > 1e-6
