"""NumPy/SciPy restatement of serinv's sequential SPD BT/BTA routines.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Every function works in place on C-contiguous float64 block arrays with the
reference layout:
    diagonal blocks          (n, b, b)
    lower-diagonal blocks    (n-1, b, b)
    lower arrow blocks       (n, a, b)
    arrow tip block          (a, a)
    right-hand side B        (n*b + a, r)   (n*b, r) for the BT solve
and follows the non-streaming reference routine step for step: the natural
ordering, or with buffer= the permuted ("middle partition") one, whose top
buffer (n, b, b) holds buffer[i] = the (0, i) block.  The Cholesky is scipy.linalg.cholesky(lower=True, overwrite_a=False,
check_finite=False), exactly the function the reference selects for NumPy
arrays (src/serinv/__init__.py:137-155, line 151); it zero-fills the upper
triangle.  Triangular solves are scipy.linalg.solve_triangular with the same
flags the reference passes.

Parity status: every function here is pinned by a dense-linear-algebra check
in tests/test_oracle_pinning.py (see the module docstring of oracle/).
"""

from functools import partial

import numpy as np
import scipy.linalg as sla

# src/serinv/__init__.py:151
_cholesky = partial(sla.cholesky, lower=True, overwrite_a=False, check_finite=False)


def _check_kwargs(kwargs, allowed):
    # src/serinv/algs/pobtaf.py:50-53, pobtasi.py:45-48, pobtf.py:44-47,
    # pobtsi.py:39-42: unknown keyword arguments raise TypeError.
    unexpected = set(kwargs) - set(allowed)
    if unexpected:
        raise TypeError(f"Unexpected keyword arguments: {unexpected}")


def _reject_unsupported(kwargs):
    # The oracle restates the in-memory branches (src/serinv/algs/pobtaf.py:
    # 58-94 and the matching branches elsewhere), not the streaming ones.
    if kwargs.get("device_streaming", False):
        raise NotImplementedError("device streaming is not restated")


# --------------------------------------------------------------------------
# BTA: pobtaf / pobtasi / pobtas
# --------------------------------------------------------------------------
def pobtaf(D, L, R, T, **kwargs):
    """Block Cholesky of an SPD BTA matrix, in place.

    Follows src/serinv/algs/pobtaf.py:11-94 (dispatch) and :97-183 (_pobtaf).
    """
    _check_kwargs(kwargs, {"device_streaming", "buffer", "factorize_last_block"})
    _reject_unsupported(kwargs)
    if kwargs.get("buffer", None) is not None:  # pobtaf.py:59-76
        return pobtaf_permuted(D, L, R, T, kwargs["buffer"])
    factorize_last_block = kwargs.get("factorize_last_block", True)
    n = D.shape[0]
    for i in range(n - 1):  # pobtaf.py:115
        D[i] = _cholesky(D[i])  # :117
        # L_{i+1,i} = A_{i+1,i} L_ii^{-T}  (:120-128)
        L[i] = sla.solve_triangular(D[i], L[i].conj().T, lower=True).conj().T
        # L_{t,i} = A_{t,i} L_ii^{-T}  (:131-139)
        R[i] = sla.solve_triangular(D[i], R[i].conj().T, lower=True).conj().T
        # Schur updates (:143-159), in the reference's order
        D[i + 1] = D[i + 1] - L[i] @ L[i].conj().T
        R[i + 1] = R[i + 1] - R[i] @ L[i].conj().T
        T[:, :] = T - R[i] @ R[i].conj().T
    if factorize_last_block:  # :161-183
        D[-1] = _cholesky(D[-1])
        R[-1] = sla.solve_triangular(D[-1], R[-1].conj().T, lower=True).conj().T
        T[:, :] = T - R[-1] @ R[-1].conj().T
        T[:, :] = _cholesky(T)


def pobtasi(D, L, R, T, **kwargs):
    """Selected inversion of an SPD BTA matrix from its block Cholesky factor.

    Follows src/serinv/algs/pobtasi.py:10-89 (dispatch) and :92-172 (_pobtasi).
    On exit D holds X_ii (full symmetric), L holds X_{i+1,i}, R holds X_{t,i}
    and T holds X_T.
    """
    _check_kwargs(kwargs, {"device_streaming", "buffer", "invert_last_block"})
    _reject_unsupported(kwargs)
    if kwargs.get("buffer", None) is not None:  # pobtasi.py:54-71
        return pobtasi_permuted(D, L, R, T, kwargs["buffer"])
    invert_last_block = kwargs.get("invert_last_block", True)
    n, b = D.shape[0], D.shape[1]
    eye_b = np.eye(b, dtype=D.dtype)  # pobtasi.py:112
    if invert_last_block:  # :114-139
        LT_inv = sla.solve_triangular(T, np.eye(T.shape[0], dtype=T.dtype), lower=True)
        T[:, :] = LT_inv.conj().T @ LT_inv  # :119
        Lt_last = R[-1].copy()  # :122
        Linv = sla.solve_triangular(D[-1], eye_b, lower=True)  # :124-128
        R[-1] = -T @ Lt_last @ Linv  # :131-133
        D[-1] = (Linv.conj().T - R[-1].conj().T @ Lt_last) @ Linv  # :136-139
    for i in range(n - 2, -1, -1):  # :141
        Li = L[i].copy()  # :142
        Lti = R[i].copy()  # :143
        Linv = sla.solve_triangular(D[i], eye_b, lower=True)  # :145-149
        L[i] = (-D[i + 1] @ Li - R[i + 1].conj().T @ Lti) @ Linv  # :153-156
        R[i] = (-R[i + 1] @ Li - T @ Lti) @ Linv  # :160-163
        D[i] = (Linv.conj().T - L[i].conj().T @ Li - R[i].conj().T @ Lti) @ Linv  # :167-172


def pobtas(D, L, R, T, B, trans="N", **kwargs):
    """Forward (trans="N") or backward (trans="T"/"C") BTA triangular solve.

    Follows src/serinv/algs/pobtas.py:10-62 (dispatch) and :65-161 (_pobtas).
    The reference does not validate kwargs for pobtas.
    """
    partial_ = kwargs.get("partial", False)
    if kwargs.get("buffer", None) is not None:  # pobtas.py:31-47
        if kwargs.get("device_streaming", False):
            raise NotImplementedError("Streaming is not implemented for the permuted arrowhead.")
        return pobtas_permuted(D, L, R, T, B, kwargs["buffer"], trans)
    if kwargs.get("device_streaming", False):
        # pobtas.py:49-52
        raise NotImplementedError("Streaming is not implemented for the natural arrowhead.")
    n, b, a = D.shape[0], D.shape[1], R.shape[1]
    blk = lambda i: slice(i * b, (i + 1) * b)  # noqa: E731
    tip = slice(n * b, n * b + a)
    if trans == "N":  # :80-126
        for i in range(n - 1):
            B[blk(i)] = sla.solve_triangular(D[i], B[blk(i)], lower=True)
            B[blk(i + 1)] -= L[i] @ B[blk(i)]
            B[tip] -= R[i] @ B[blk(i)]
        if not partial_:
            B[blk(n - 1)] = sla.solve_triangular(D[n - 1], B[blk(n - 1)], lower=True)
            B[tip] -= R[-1] @ B[blk(n - 1)]
            B[tip] = sla.solve_triangular(T, B[tip], lower=True)
    elif trans == "T" or trans == "C":  # :127-159
        if not partial_:
            B[tip] = sla.solve_triangular(T, B[tip], lower=True, trans="C")
            B[blk(n - 1)] = sla.solve_triangular(
                D[-1], B[blk(n - 1)] - R[-1].conj().T @ B[tip], lower=True, trans="C"
            )
        for i in range(n - 2, -1, -1):
            B[blk(i)] = sla.solve_triangular(
                D[i],
                B[blk(i)] - L[i].conj().T @ B[blk(i + 1)] - R[i].conj().T @ B[tip],
                lower=True,
                trans="C",
            )
    else:
        raise ValueError(f"Invalid transpose argument: {trans}.")  # :160-161


# --------------------------------------------------------------------------
# Permuted ("middle partition") forms: blocks 1 .. n-2 are eliminated; block
# 0, block n-1 and the tip form the reduced system.  a == 0 (R of shape
# (n, 0, b), T of shape (0, 0)) gives the BT forms pobtf / pobtsi / pobts
# with buffer=, whose steps are the same without the arrow terms.
# --------------------------------------------------------------------------
def pobtaf_permuted(D, L, R, T, buffer):
    """Follows src/serinv/algs/pobtaf.py:186-281 (_pobtaf_permuted) and, for
    a == 0, pobtf.py:125-185 (_pobtf_permuted)."""
    n = D.shape[0]
    buffer[1] = L[0].conj().T  # pobtaf.py:204
    for i in range(1, n - 1):  # :207
        D[i] = _cholesky(D[i])  # :209
        L[i] = sla.solve_triangular(D[i], L[i].conj().T, lower=True).conj().T  # :213-221
        buffer[i] = sla.solve_triangular(D[i], buffer[i].conj().T, lower=True).conj().T  # :224-232
        R[i] = sla.solve_triangular(D[i], R[i].conj().T, lower=True).conj().T  # :235-243
        D[i + 1] = D[i + 1] - L[i] @ L[i].conj().T  # :247-251
        R[i + 1] = R[i + 1] - R[i] @ L[i].conj().T  # :254-257
        T[:, :] = T - R[i] @ R[i].conj().T  # :261-264
        D[0] = D[0] - buffer[i] @ buffer[i].conj().T  # :268-270
        buffer[i + 1] = -buffer[i] @ L[i].conj().T  # :273-275
        R[0] = R[0] - R[i] @ buffer[i].conj().T  # :279-281


def pobtasi_permuted(D, L, R, T, buffer):
    """Follows src/serinv/algs/pobtasi.py:175-244 (_pobtasi_permuted) and, for
    a == 0, pobtsi.py:128-179 (_pobtsi_permuted)."""
    n, b = D.shape[0], D.shape[1]
    eye_b = np.eye(b, dtype=D.dtype)
    for i in range(n - 2, 0, -1):  # pobtasi.py:199
        Li = L[i].copy()  # :200
        Ri = R[i].copy()  # :201
        Bi = buffer[i].copy()  # :202
        Linv = sla.solve_triangular(D[i], eye_b, lower=True)  # :204-208
        L[i] = (-buffer[i + 1].conj().T @ Bi - D[i + 1] @ Li - R[i + 1].conj().T @ Ri) @ Linv  # :211-217
        buffer[i] = (-buffer[i + 1] @ Li - D[0] @ Bi - R[0].conj().T @ Ri) @ Linv  # :220-224
        R[i] = (-R[i + 1] @ Li - R[0] @ Bi - T @ Ri) @ Linv  # :228-232
        D[i] = (Linv.conj().T - L[i].conj().T @ Li - buffer[i].conj().T @ Bi
                - R[i].conj().T @ Ri) @ Linv  # :235-241
    L[0] = buffer[1].conj().T  # :244


def pobtas_permuted(D, L, R, T, B, buffer, trans):
    """Follows src/serinv/algs/pobtas.py:164-218 (_pobtas_permuted) and, for
    a == 0, pobts.py:121-165 (_pobts_permuted)."""
    n, b, a = D.shape[0], D.shape[1], R.shape[1]
    blk = lambda i: slice(i * b, (i + 1) * b)  # noqa: E731
    tip = slice(n * b, n * b + a)
    if trans == "N":  # pobtas.py:179-204
        for i in range(1, n - 1):
            B[blk(i)] = sla.solve_triangular(D[i], B[blk(i)], lower=True)
            B[blk(i + 1)] -= L[i] @ B[blk(i)]
            B[blk(0)] -= buffer[i] @ B[blk(i)]
            B[tip] -= R[i] @ B[blk(i)]
    elif trans == "T" or trans == "C":  # :205-216
        for i in range(n - 2, 0, -1):
            B[blk(i)] = sla.solve_triangular(
                D[i],
                B[blk(i)] - L[i].conj().T @ B[blk(i + 1)] - R[i].conj().T @ B[tip]
                - buffer[i].conj().T @ B[blk(0)],
                lower=True,
                trans="C",
            )
    else:
        raise ValueError(f"Invalid transpose argument: {trans}.")  # :217-218


# --------------------------------------------------------------------------
# BT (arrowless): pobtf / pobtsi / pobts
# --------------------------------------------------------------------------
def _no_arrow(D):
    return np.zeros((D.shape[0], 0, D.shape[1]), dtype=D.dtype)


def pobtf(D, L, **kwargs):
    """Block Cholesky of an SPD BT matrix (src/serinv/algs/pobtf.py:11-122)."""
    _check_kwargs(kwargs, {"device_streaming", "buffer", "factorize_last_block"})
    _reject_unsupported(kwargs)
    if kwargs.get("buffer", None) is not None:  # pobtf.py:53-68
        return pobtaf_permuted(D, L, _no_arrow(D), np.zeros((0, 0)), kwargs["buffer"])
    factorize_last_block = kwargs.get("factorize_last_block", True)
    n = D.shape[0]
    for i in range(n - 1):  # pobtf.py:97
        D[i] = _cholesky(D[i])  # :99
        L[i] = sla.solve_triangular(D[i], L[i].conj().T, lower=True).conj().T  # :102-110
        D[i + 1] = D[i + 1] - L[i] @ L[i].conj().T  # :114-118
    if factorize_last_block:
        D[-1] = _cholesky(D[-1])  # :122


def pobtsi(D, L, **kwargs):
    """Selected inversion of an SPD BT matrix (src/serinv/algs/pobtsi.py:10-125)."""
    _check_kwargs(kwargs, {"device_streaming", "buffer", "invert_last_block"})
    _reject_unsupported(kwargs)
    if kwargs.get("buffer", None) is not None:  # pobtsi.py:48-62
        return pobtasi_permuted(D, L, _no_arrow(D), np.zeros((0, 0)), kwargs["buffer"])
    invert_last_block = kwargs.get("invert_last_block", True)
    n, b = D.shape[0], D.shape[1]
    eye_b = np.eye(b, dtype=D.dtype)  # pobtsi.py:92
    if invert_last_block:  # :94-102
        Linv = sla.solve_triangular(D[-1], eye_b, lower=True)
        D[-1] = Linv.conj().T @ Linv
    for i in range(n - 2, -1, -1):  # :104
        Li = L[i].copy()  # :105
        Linv = sla.solve_triangular(D[i], eye_b, lower=True)  # :107-111
        L[i] = (-D[i + 1] @ Li) @ Linv  # :115-117
        D[i] = (Linv.conj().T - L[i].conj().T @ Li) @ Linv  # :121-125


def pobts(D, L, B, trans="N", **kwargs):
    """Forward/backward BT triangular solve (src/serinv/algs/pobts.py:10-118)."""
    partial_ = kwargs.get("partial", False)
    if kwargs.get("buffer", None) is not None:
        if kwargs.get("device_streaming", False):  # pobts.py:29-40
            raise NotImplementedError("Permuted arrowhead is not implemented.")
        return pobtas_permuted(D, L, _no_arrow(D), np.zeros((0, 0)), B, kwargs["buffer"], trans)
    if kwargs.get("device_streaming", False):
        raise NotImplementedError("Streaming is not implemented for the natural arrowhead.")
    n, b = D.shape[0], D.shape[1]
    blk = lambda i: slice(i * b, (i + 1) * b)  # noqa: E731
    if trans == "N":  # pobts.py:69-95
        for i in range(n - 1):
            B[blk(i)] = sla.solve_triangular(D[i], B[blk(i)], lower=True)
            B[blk(i + 1)] -= L[i] @ B[blk(i)]
        if not partial_:
            B[blk(n - 1)] = sla.solve_triangular(D[n - 1], B[blk(n - 1)], lower=True)
    elif trans == "T" or trans == "C":  # :96-116
        if not partial_:
            B[blk(n - 1)] = sla.solve_triangular(D[-1], B[blk(n - 1)], lower=True, trans="C")
        for i in range(n - 2, -1, -1):
            B[blk(i)] = sla.solve_triangular(
                D[i], B[blk(i)] - L[i].conj().T @ B[blk(i + 1)], lower=True, trans="C"
            )
    else:
        raise ValueError(f"Invalid transpose argument: {trans}.")  # :117-118


# --------------------------------------------------------------------------
# Build-side extension (no reference symbol): log-determinant from the factor
# --------------------------------------------------------------------------
def logdet_from_factor(D, T=None):
    """log det A = 2 * sum(log diag L_ii) + 2 * sum(log diag L_T).

    D holds the pobtaf/pobtf output (lower Cholesky blocks); T the tip factor
    (None or empty for BT).  Pinned by numpy.linalg.slogdet.
    """
    s = 2.0 * np.sum(np.log(np.diagonal(D, axis1=1, axis2=2)))
    if T is not None and T.size:
        s += 2.0 * np.sum(np.log(np.diag(T)))
    return float(s)


# --------------------------------------------------------------------------
# Non-SPD (diagonally dominant) family: ddbtasc / ddbtasci, ddbtsc / ddbtsci.
# Block arrays: diagonal (n,b,b), lower (n-1,b,b) = A_{i+1,i}, upper
# (n-1,b,b) = A_{i,i+1}, lower arrow (n,a,b) = A_{t,i}, upper arrow (n,b,a)
# = A_{i,t}, tip (a,a).  The block inverses are numpy.linalg.inv, which the
# reference calls (LU with partial pivoting).  Only the "AX=I" forms
# (no rhs, no buffers) are restated.
# --------------------------------------------------------------------------
def ddbtasc(D, Lo, Up, Rl, Ru, T, invert_last_block=True):
    """Schur-complement (block LU) sweep, src/serinv/algs/ddbtasc.py:187-241."""
    n = D.shape[0]
    for i in range(1, n):  # :199
        D[i - 1] = np.linalg.inv(D[i - 1])  # :201
        D[i] = D[i] - Lo[i - 1] @ D[i - 1] @ Up[i - 1]  # :204-209
        Rl[i] = Rl[i] - Rl[i - 1] @ D[i - 1] @ Up[i - 1]  # :212-217
        Ru[i] = Ru[i] - Lo[i - 1] @ D[i - 1] @ Ru[i - 1]  # :220-225
        T[:, :] = T - Rl[i - 1] @ D[i - 1] @ Ru[i - 1]  # :228-233
    if invert_last_block:  # :235-241
        D[-1] = np.linalg.inv(D[-1])
        T[:, :] = np.linalg.inv(T - Rl[-1] @ D[-1] @ Ru[-1])


def ddbtasci(D, Lo, Up, Rl, Ru, T, invert_last_block=True):
    """Selected inversion from the ddbtasc output, src/serinv/algs/ddbtasci.py:187-256."""
    n = D.shape[0]
    if invert_last_block:  # :207-219
        B2 = D[-1] @ Ru[-1]
        Rl[-1] = -T @ Rl[-1] @ D[-1]
        Ru[-1] = -B2 @ T
        D[-1] = D[-1] - B2 @ Rl[-1]
    for i in range(n - 2, -1, -1):  # :221
        B1 = Up[i] @ D[i + 1] + Ru[i] @ Rl[i + 1]  # :222-225
        B2 = Up[i] @ Ru[i + 1] + Ru[i] @ T  # :227-230
        C1 = D[i + 1] @ Lo[i] + Ru[i + 1] @ Rl[i]  # :232-235
        C2 = Rl[i + 1] @ Lo[i] + T @ Rl[i]  # :237-240
        Up[i] = -D[i] @ B1  # :242
        Ru[i] = -D[i] @ B2  # :243
        D1 = Lo[i].copy()  # :245
        D2 = Rl[i].copy()  # :246
        Lo[i] = -C1 @ D[i]  # :248
        Rl[i] = -C2 @ D[i]  # :249
        D[i] = D[i] + D[i] @ (B1 @ D1 + B2 @ D2) @ D[i]  # :251-256


def ddbtsc(D, Lo, Up, invert_last_block=True, direction="downward"):
    """src/serinv/algs/ddbtsc.py:180-221 (downward and upward)."""
    n = D.shape[0]
    if direction == "downward":
        for i in range(0, n - 1):  # :188
            D[i] = np.linalg.inv(D[i])
            D[i + 1] = D[i + 1] - Lo[i] @ D[i] @ Up[i]
        if invert_last_block:
            D[-1] = np.linalg.inv(D[-1])
    elif direction == "upward":
        for i in range(n - 1, 0, -1):  # :210
            D[i] = np.linalg.inv(D[i])
            D[i - 1] = D[i - 1] - Up[i - 1] @ D[i] @ Lo[i - 1]
        if invert_last_block:
            D[0] = np.linalg.inv(D[0])
    else:
        raise ValueError("Optional keyword argument `direction` must be either 'downward' or 'upward'")


def ddbtsci(D, Lo, Up, direction="downward"):
    """src/serinv/algs/ddbtsci.py:176-237 (downward and upward)."""
    n = D.shape[0]
    if direction == "downward":
        for i in range(n - 2, -1, -1):  # :188
            tl = Lo[i].copy()
            Lo[i] = -D[i + 1] @ Lo[i] @ D[i]
            Up[i] = -D[i] @ Up[i] @ D[i + 1]
            D[i] = D[i] - Up[i] @ tl @ D[i]
    elif direction == "upward":
        for i in range(1, n):  # :220
            tu = Up[i - 1].copy()
            Lo[i - 1] = -D[i] @ Lo[i - 1] @ D[i - 1]
            Up[i - 1] = -D[i - 1] @ Up[i - 1] @ D[i]
            D[i] = D[i] - Lo[i - 1] @ tu @ D[i]
    else:
        raise ValueError("Optional keyword argument `direction` must be either 'downward' or 'upward'")


def ddbtsc_quadratic(D, Lo, Up, BD, BLo, BUp, invert_last_block=True, direction="downward"):
    """A X A^T = B sweep of a BT matrix: downward src/serinv/algs/ddbtsc.py:271-311,
    upward :314-356.  Each step inverts the block being eliminated, maps its
    B block to A_ii^{-1} B_ii A_ii^{-T}, and updates the neighbour's A and B
    blocks (loops ddbtsc.py:285-305 downward, :328-348 upward)."""
    n = D.shape[0]
    if direction == "downward":
        for i in range(0, n - 1):
            D[i] = np.linalg.inv(D[i])
            BD[i] = D[i] @ BD[i] @ D[i].T
            LD = Lo[i] @ D[i]  # A_{i+1,i} A_ii^{-1}
            D[i + 1] = D[i + 1] - LD @ Up[i]
            BD[i + 1] = BD[i + 1] + Lo[i] @ BD[i] @ Lo[i].T - BLo[i] @ LD.T - LD @ BUp[i]
        last = n - 1
    elif direction == "upward":
        for i in range(n - 1, 0, -1):
            D[i] = np.linalg.inv(D[i])
            BD[i] = D[i] @ BD[i] @ D[i].T
            UD = Up[i - 1] @ D[i]  # A_{i-1,i} A_ii^{-1}
            D[i - 1] = D[i - 1] - UD @ Lo[i - 1]
            BD[i - 1] = BD[i - 1] + Up[i - 1] @ BD[i] @ Up[i - 1].T - BUp[i - 1] @ UD.T - UD @ BLo[i - 1]
        last = 0
    else:
        raise ValueError("Optional keyword argument `direction` must be either 'downward' or 'upward'")
    if invert_last_block:  # :307-311, :350-356
        D[last] = np.linalg.inv(D[last])
        BD[last] = D[last] @ BD[last] @ D[last].T


def ddbtsci_quadratic(D, Lo, Up, BD, BLo, BUp, direction="downward"):
    """Selected inversion for A X A^T = B of a BT matrix from ddbtsc_quadratic's
    output (same direction): downward src/serinv/algs/ddbtsci.py:299-369
    (loop :320), upward :372-445 (loop :393).  Each step pairs the finished
    block f (X_ff, XB_ff) with its neighbour g, which holds the Schur inverse
    S_g^{-1} and S_g^{-1} B~_g S_g^{-T}.  In g/f roles both directions run
    the same formulas: P = S_g^{-1} A_gf (the reference's temp_1 downward,
    temp_3 upward) and Q = X_ff A_fg (temp_3 downward, temp_1 upward).  For a
    symmetric B (the reference uses B_gg^T for B_gg)."""
    n = D.shape[0]
    if direction == "downward":
        # g = i, f = i + 1; A_gf = Up[i], A_fg = Lo[i]
        steps = [(i, i + 1, i) for i in range(n - 2, -1, -1)]
        AU, AL, BU, BL = Up, Lo, BUp, BLo
    elif direction == "upward":
        # g = i, f = i - 1; A_gf = Lo[i-1], A_fg = Up[i-1]
        steps = [(i, i - 1, i - 1) for i in range(1, n)]
        AU, AL, BU, BL = Lo, Up, BLo, BUp
    else:
        raise ValueError("Optional keyword argument `direction` must be either 'downward' or 'upward'")
    for g, f, k in steps:
        Bgf, Bfg = BU[k].copy(), BL[k].copy()
        t1 = D[g] @ AU[k]         # P
        t4 = AL[k].T @ D[f].T     # Q^T
        t2 = AU[k].T @ D[g].T     # P^T
        t3 = D[f] @ AL[k]         # Q
        BU[k] = -t1 @ BD[f] - BD[g] @ t4 + D[g] @ Bgf @ D[f].T
        BL[k] = -BD[f] @ t2 - t3 @ BD[g] + D[f] @ Bfg @ D[g].T
        BD[g] = (BD[g] + t1 @ BD[f] @ t2 + t1 @ t3 @ BD[g] + BD[g].T @ t4 @ t2
                 - t1 @ D[f] @ Bfg @ D[g].T - D[g] @ Bgf @ D[f].T @ t2)
        Afg = AL[k].copy()
        AL[k] = -t3 @ D[g]
        AU[k] = -t1 @ D[f]
        D[g] = D[g] - AU[k] @ Afg @ D[g]


def ddbtasc_quadratic(D, Lo, Up, Rl, Ru, T, BD, BLo, BUp, BRl, BRu, BT, invert_last_block=True):
    """Schur-complement sweep for A X A^T = B, src/serinv/algs/ddbtasc.py:335-475:
    the A sweep of ddbtasc, with B's blocks carried through the same
    elimination (B[i-1] <- A_ii^{-1} B[i-1] A_ii^{-T} and the updates of the
    next block row)."""
    n = D.shape[0]
    for i in range(1, n):  # :351
        D[i - 1] = np.linalg.inv(D[i - 1])  # :353
        D[i] = D[i] - Lo[i - 1] @ D[i - 1] @ Up[i - 1]  # :356-361
        Rl[i] = Rl[i] - Rl[i - 1] @ D[i - 1] @ Up[i - 1]  # :364-369
        Ru[i] = Ru[i] - Lo[i - 1] @ D[i - 1] @ Ru[i - 1]  # :372-377
        T[:, :] = T - Rl[i - 1] @ D[i - 1] @ Ru[i - 1]  # :380-385
        Di, L, R = D[i - 1], Lo[i - 1], Rl[i - 1]
        BD[i - 1] = Di @ BD[i - 1] @ Di.T  # :389-393
        BD[i] = BD[i] + L @ BD[i - 1] @ L.T - BLo[i - 1] @ Di.T @ L.T - L @ Di @ BUp[i - 1]  # :395-406
        BRu[i] = BRu[i] + L @ BD[i - 1] @ R.T - BLo[i - 1] @ Di.T @ R.T - L @ Di @ BRu[i - 1]  # :408-419
        BRl[i] = BRl[i] + R @ BD[i - 1] @ L.T - BRl[i - 1] @ Di.T @ L.T - R @ Di @ BUp[i - 1]  # :421-432
        BT[:, :] = BT + R @ BD[i - 1] @ R.T - BRl[i - 1] @ Di.T @ R.T - R @ Di @ BRu[i - 1]  # :434-445
    if invert_last_block:  # :447-475
        D[-1] = np.linalg.inv(D[-1])
        BD[-1] = D[-1] @ BD[-1] @ D[-1].T
        T[:, :] = np.linalg.inv(T - Rl[-1] @ D[-1] @ Ru[-1])
        BT[:, :] = T @ (BT + Rl[-1] @ BD[-1] @ Rl[-1].T - BRl[-1] @ D[-1].T @ Rl[-1].T
                        - Rl[-1] @ D[-1] @ BRu[-1]) @ T.T


def ddbtasci_quadratic(D, Lo, Up, Rl, Ru, T, BD, BLo, BUp, BRl, BRu, BT, invert_last_block=True):
    """Selected inversion for A X A^T = B from ddbtasc_quadratic's output,
    src/serinv/algs/ddbtasci.py:343-576: on exit D.. hold the blocks of
    A^{-1} and BD.. those of A^{-1} B A^{-T}.  Like the reference, for a
    symmetric B (it uses B_ii^T for B_ii, :540-544)."""
    n = D.shape[0]
    if invert_last_block:  # :377-430
        B2 = D[-1] @ Ru[-1]
        C2 = Ru[-1].T @ D[-1].T
        D2 = T @ Rl[-1] @ BD[-1]
        t13 = BRu[-1].copy()
        t31 = BRl[-1].copy()
        BRu[-1] = -B2 @ BT - BD[-1] @ Rl[-1].T @ T.T + D[-1] @ BRu[-1] @ T.T
        BRl[-1] = -BT @ C2 - D2 + T @ BRl[-1] @ D[-1].T
        BD[-1] = (BD[-1] + B2 @ BT @ C2 + B2 @ D2 + BD[-1].T @ Rl[-1].T @ T.T @ C2
                  - B2 @ T @ t31 @ D[-1].T - D[-1] @ t13 @ T.T @ C2)
        Rl[-1] = -T @ Rl[-1] @ D[-1]
        Ru[-1] = -B2 @ T
        D[-1] = D[-1] - B2 @ Rl[-1]
    for i in range(n - 2, -1, -1):  # :432
        B1 = Up[i] @ D[i + 1] + Ru[i] @ Rl[i + 1]
        B2 = Up[i] @ Ru[i + 1] + Ru[i] @ T
        C1 = D[i + 1] @ Lo[i] + Ru[i + 1] @ Rl[i]
        C2 = Rl[i + 1] @ Lo[i] + T @ Rl[i]
        t12, t13, t21, t31 = BUp[i].copy(), BRu[i].copy(), BLo[i].copy(), BRl[i].copy()  # :453-456
        BUp[i] = (-D[i] @ (Up[i] @ BD[i + 1] + Ru[i] @ BRl[i + 1]) - BD[i] @ C1.T
                  + D[i] @ (BUp[i] @ D[i + 1].T + BRu[i] @ Ru[i + 1].T))  # :458-469
        BRu[i] = (-D[i] @ (Up[i] @ BRu[i + 1] + Ru[i] @ BT) - BD[i] @ C2.T
                  + D[i] @ (t12 @ Rl[i + 1].T + BRu[i] @ T.T))  # :470-483
        BLo[i] = (-(BD[i + 1] @ Up[i].T + BRu[i + 1] @ Ru[i].T) @ D[i].T - C1 @ BD[i]
                  + (D[i + 1] @ BLo[i] + Ru[i + 1] @ BRl[i]) @ D[i].T)  # :485-496
        BRl[i] = (-(BRl[i + 1] @ Up[i].T + BT @ Ru[i].T) @ D[i].T - C2 @ BD[i]
                  + (Rl[i + 1] @ t21 + T @ BRl[i]) @ D[i].T)  # :497-508
        BD[i] = (BD[i]
                 + D[i] @ ((Up[i] @ BD[i + 1] + Ru[i] @ BRl[i + 1]) @ Up[i].T
                           + (Up[i] @ BRu[i + 1] + Ru[i] @ BT) @ Ru[i].T) @ D[i].T
                 + D[i] @ (B1 @ Lo[i] + B2 @ Rl[i]) @ BD[i]
                 + BD[i].T @ (C1.T @ Up[i].T + C2.T @ Ru[i].T) @ D[i].T
                 - D[i] @ (B1 @ t21 + B2 @ t31) @ D[i].T
                 - D[i] @ ((t12 @ D[i + 1].T + t13 @ Ru[i + 1].T) @ Up[i].T
                           + (t12 @ Rl[i + 1].T + t13 @ T.T) @ Ru[i].T) @ D[i].T)  # :510-559
        Up[i] = -D[i] @ B1  # :562
        Ru[i] = -D[i] @ B2
        D1 = Lo[i].copy()
        D2 = Rl[i].copy()
        Lo[i] = -C1 @ D[i]
        Rl[i] = -C2 @ D[i]
        D[i] = D[i] + D[i] @ (B1 @ D1 + B2 @ D2) @ D[i]  # :571-576


# --------------------------------------------------------------------------
# Non-SPD permuted ("middle partition") forms.  Lbuf, Ubuf have the shape of
# the lower blocks, (n-1, b, b): during the sweep Lbuf[k] holds the fill-in
# A~_{0,k+1} of row 0 and Ubuf[k] the fill-in A~_{k+1,0} of column 0; after
# the selected inversion they hold X_{0,k+1} and X_{k+1,0}.  The middle
# blocks 1 .. n-2 are eliminated; blocks 0, n-1 and the tip form the reduced
# system.
# --------------------------------------------------------------------------
def ddbtasc_permuted(D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf):
    """src/serinv/algs/ddbtasc.py:244-332 (with a = 0, ddbtsc.py:224-269)."""
    n = D.shape[0]
    Lbuf[0] = Up[0]  # A_{0,1}
    Ubuf[0] = Lo[0]  # A_{1,0}
    for i in range(1, n - 1):
        D[i] = np.linalg.inv(D[i])
        S = D[i]
        D[i + 1] = D[i + 1] - Lo[i] @ S @ Up[i]
        Rl[i + 1] = Rl[i + 1] - Rl[i] @ S @ Up[i]
        Ru[i + 1] = Ru[i + 1] - Lo[i] @ S @ Ru[i]
        T[:, :] = T - Rl[i] @ S @ Ru[i]
        Lbuf[i] = -Lbuf[i - 1] @ S @ Up[i]
        Ubuf[i] = -Lo[i] @ S @ Ubuf[i - 1]
        D[0] = D[0] - Lbuf[i - 1] @ S @ Ubuf[i - 1]
        Rl[0] = Rl[0] - Rl[i] @ S @ Ubuf[i - 1]
        Ru[0] = Ru[0] - Lbuf[i - 1] @ S @ Ru[i]


def ddbtasci_permuted(D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf):
    """src/serinv/algs/ddbtasci.py:259-341 (with a = 0, ddbtsci.py:240-297).
    Expects blocks 0, n-1, the tip, their arrow blocks and Lbuf[-1], Ubuf[-1]
    to hold the reduced system's inverse."""
    n = D.shape[0]
    for i in range(n - 2, 0, -1):
        # row i of A~ times the finished columns, and the finished rows times column i
        B1 = Up[i] @ D[i + 1] + Ru[i] @ Rl[i + 1] + Ubuf[i - 1] @ Lbuf[i]
        B2 = Up[i] @ Ubuf[i] + Ubuf[i - 1] @ D[0] + Ru[i] @ Rl[0]
        B3 = Up[i] @ Ru[i + 1] + Ru[i] @ T + Ubuf[i - 1] @ Ru[0]
        C1 = D[i + 1] @ Lo[i] + Ru[i + 1] @ Rl[i] + Ubuf[i] @ Lbuf[i - 1]
        C2 = Lbuf[i] @ Lo[i] + D[0] @ Lbuf[i - 1] + Ru[0] @ Rl[i]
        C3 = Rl[i + 1] @ Lo[i] + T @ Rl[i] + Rl[0] @ Lbuf[i - 1]
        S = D[i]
        E = B1 @ Lo[i] + B2 @ Lbuf[i - 1] + B3 @ Rl[i]
        Up[i] = -S @ B1
        Ubuf[i - 1] = -S @ B2
        Ru[i] = -S @ B3
        Lo[i] = -C1 @ S
        Lbuf[i - 1] = -C2 @ S
        Rl[i] = -C3 @ S
        D[i] = S + S @ E @ S
    if n > 1:
        Lo[0] = Ubuf[0]  # X_{1,0}
        Up[0] = Lbuf[0]  # X_{0,1}


def ddbtasc_quadratic_permuted(D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf, BD, BLo, BUp, BRl, BRu, BT,
                               BLbuf, BUbuf):
    """Permuted sweep for A X A^T = B, src/serinv/algs/ddbtasc.py:478-710 (with
    a = 0, ddbtsc.py:357-466): the A sweep of ddbtasc_permuted, then B's blocks
    carried through the same elimination.  With F_r = A~_{r,i}, G_r =
    B~_{r,i}, H_r = B~_{i,r} for the neighbours r in (i+1, t, 0) of the
    eliminated block i, and S = A~_ii^{-1}: B~_ii <- S B~_ii S^T, then
    B~_rc += F_r B~_ii F_c^T - G_r S^T F_c^T - F_r S H_c for every pair
    (ddbtasc.py:581-708).  BLbuf[k], BUbuf[k] for k >= 1 accumulate, so they
    must start at zero (the reference's allocator zeroes them)."""
    n = D.shape[0]
    Lbuf[0] = Up[0]  # :498-501
    Ubuf[0] = Lo[0]
    BLbuf[0] = BUp[0]
    BUbuf[0] = BLo[0]
    for i in range(1, n - 1):
        D[i] = np.linalg.inv(D[i])
        S = D[i]
        D[i + 1] = D[i + 1] - Lo[i] @ S @ Up[i]
        Rl[i + 1] = Rl[i + 1] - Rl[i] @ S @ Up[i]
        Ru[i + 1] = Ru[i + 1] - Lo[i] @ S @ Ru[i]
        T[:, :] = T - Rl[i] @ S @ Ru[i]
        Lbuf[i] = -Lbuf[i - 1] @ S @ Up[i]
        Ubuf[i] = -Lo[i] @ S @ Ubuf[i - 1]
        D[0] = D[0] - Lbuf[i - 1] @ S @ Ubuf[i - 1]
        Rl[0] = Rl[0] - Rl[i] @ S @ Ubuf[i - 1]
        Ru[0] = Ru[0] - Lbuf[i - 1] @ S @ Ru[i]
        BD[i] = S @ BD[i] @ S.T  # :581-583
        F = {"n": Lo[i], "t": Rl[i], "0": Lbuf[i - 1]}
        G = {"n": BLo[i], "t": BRl[i], "0": BLbuf[i - 1]}
        H = {"n": BUp[i], "t": BRu[i], "0": BUbuf[i - 1]}

        def upd(r, c):
            return F[r] @ BD[i] @ F[c].T - G[r] @ S.T @ F[c].T - F[r] @ S @ H[c]

        BD[i + 1] = BD[i + 1] + upd("n", "n")
        BRl[i + 1] = BRl[i + 1] + upd("t", "n")
        BRu[i + 1] = BRu[i + 1] + upd("n", "t")
        BT[:, :] = BT + upd("t", "t")
        BLbuf[i] = BLbuf[i] + upd("0", "n")
        BUbuf[i] = BUbuf[i] + upd("n", "0")
        BD[0] = BD[0] + upd("0", "0")
        BRl[0] = BRl[0] + upd("t", "0")
        BRu[0] = BRu[0] + upd("0", "t")


def ddbtasci_quadratic_permuted(D, Lo, Up, Rl, Ru, T, Lbuf, Ubuf, BD, BLo, BUp, BRl, BRu, BT,
                                BLbuf, BUbuf):
    """Permuted selected inversion for A X A^T = B, src/serinv/algs/ddbtasci.py:
    578-845 (with a = 0, ddbtsci.py:446-639).  Expects the reduced system's
    inverse (A's and B's) in blocks 0, n-1, the tip, their arrow blocks and
    the last buffer blocks.  Step i (n-2 .. 1) works with the finished blocks
    j, k in (i+1, 0, t): X_jk and XB_jk (A^{-1} and A^{-1} B A^{-T}), A~'s row
    i (a_j = A~_{i,j}) and column i (c_j = A~_{j,i}), and B~'s (bR_j =
    B~_{i,j}, bC_j = B~_{j,i}); the reference's B1..B3 / C1..C3 are B_j =
    sum_k a_k X_kj and C_j = sum_k X_jk c_k (:619-653).  For a symmetric B."""
    n = D.shape[0]
    for i in range(n - 2, 0, -1):
        S, SB = D[i].copy(), BD[i].copy()
        a = [Up[i], Ubuf[i - 1], Ru[i]]
        c = [Lo[i], Lbuf[i - 1], Rl[i]]
        X = [[D[i + 1], Ubuf[i], Ru[i + 1]], [Lbuf[i], D[0], Ru[0]], [Rl[i + 1], Rl[0], T]]
        XB = [[BD[i + 1], BUbuf[i], BRu[i + 1]], [BLbuf[i], BD[0], BRu[0]],
              [BRl[i + 1], BRl[0], BT]]
        bR = [BUp[i].copy(), BUbuf[i - 1].copy(), BRu[i].copy()]
        bC = [BLo[i].copy(), BLbuf[i - 1].copy(), BRl[i].copy()]
        J = range(3)
        Bj = [sum(a[k] @ X[k][j] for k in J) for j in J]
        Cj = [sum(X[j][k] @ c[k] for k in J) for j in J]
        DF = [sum(a[k] @ XB[k][j] for k in J) for j in J]  # :664-709
        Gj = [sum(bR[k] @ X[j][k].T for k in J) for j in J]
        DK = [sum(XB[j][k] @ a[k].T for k in J) for j in J]  # :710-755
        Hj = [sum(X[j][k] @ bC[k] for k in J) for j in J]
        newR = [-S @ DF[j] - SB @ Cj[j].T + S @ Gj[j] for j in J]
        newC = [-DK[j] @ S.T - Cj[j] @ SB + Hj[j] @ S.T for j in J]
        E = sum(Bj[j] @ c[j] for j in J)
        newBD = (SB + S @ sum(DF[j] @ a[j].T for j in J) @ S.T + S @ E @ SB
                 + SB.T @ sum(Cj[j].T @ a[j].T for j in J) @ S.T
                 - S @ sum(Bj[j] @ bC[j] for j in J) @ S.T
                 - S @ sum(Gj[j] @ a[j].T for j in J) @ S.T)  # :756-821
        BUp[i], BUbuf[i - 1], BRu[i] = newR
        BLo[i], BLbuf[i - 1], BRl[i] = newC
        BD[i] = newBD
        Up[i], Ubuf[i - 1], Ru[i] = (-S @ Bj[j] for j in J)  # :822-839
        Lo[i], Lbuf[i - 1], Rl[i] = (-Cj[j] @ S for j in J)
        D[i] = S + S @ E @ S
    if n > 1:  # :841-845
        Lo[0] = Ubuf[0]
        Up[0] = Lbuf[0]
        BLo[0] = BUbuf[0]
        BUp[0] = BLbuf[0]
