"""Mixed FP32/FP64 form of the SPD BTA path (no counterpart in the reference).

    pobtaf_mixed(D, L, R, T)            # factor: FP32 tile products, FP64 potrf
    logdet = pobta_logdet(D, T)          # FP64 sum over the FP64 diagonal factors
    pobtas_refine(A, (D, L, R, T), B)    # FP64 iterative refinement of A x = B
    pobtasi_mixed(D, L, R, T)            # selected inversion, FP32 products

The trsm / syrk / gemm tile products round their FP64 operands to FP32 and
accumulate in FP32 MFMA (rocbta_mpobtaf / rocbta_mpobtasi); the blocks stay
FP64 in memory, and the potrf / trtri of every running Schur-complement
diagonal tile, the arrow-tip factorization and the log-determinant sum run
in FP64.  The factor, the selected inverse and a single solve therefore carry
errors of order cond(A) * 1e-7; pobtas_refine wins FP64 accuracy back for
solves by iterative refinement: x <- x + A^{-1}_mixed (B - A x) with the
residual formed in FP64 from the original blocks (rocbta_dresidual), as
long as cond(A) * 1e-7 < 1.
"""

import torch

from . import _lib
from .algs import pobtas
from .algs._common import (_tensor, bta_shapes, check_kwargs, ptr, reject_unsupported, stream_of,
                           workspace)
from .algs.pobtaf import _factor
from .algs.pobtasi import _selinv


def pobtaf_mixed(A_diagonal_blocks, A_lower_diagonal_blocks, A_lower_arrow_blocks,
                 A_arrow_tip_block, **kwargs):
    """pobtaf (same arguments and results) with FP32 tile products."""
    check_kwargs(kwargs, {"device_streaming", "buffer", "factorize_last_block"})
    reject_unsupported(kwargs, "pobtaf_mixed")
    D, L, R, T = A_diagonal_blocks, A_lower_diagonal_blocks, A_lower_arrow_blocks, A_arrow_tip_block
    n, b, a, dev = bta_shapes(D, L, R, T)
    _factor(D, L, R, T, n, b, a, dev, bool(kwargs.get("factorize_last_block", True)),
            "pobtaf_mixed", mixed=True)


def pobtasi_mixed(L_diagonal_blocks, L_lower_diagonal_blocks, L_lower_arrow_blocks,
                  L_arrow_tip_block, **kwargs):
    """pobtasi (same arguments and results) with FP32 tile products."""
    check_kwargs(kwargs, {"device_streaming", "buffer", "invert_last_block"})
    reject_unsupported(kwargs, "pobtasi_mixed")
    D, L, R, T = L_diagonal_blocks, L_lower_diagonal_blocks, L_lower_arrow_blocks, L_arrow_tip_block
    n, b, a, dev = bta_shapes(D, L, R, T)
    _selinv(D, L, R, T, n, b, a, dev, bool(kwargs.get("invert_last_block", True)),
            "pobtasi_mixed", mixed=True)


def residual(A, X, B, out=None):
    """out = B - A X in FP64 (A: the original blocks (D, L, R, T))."""
    D, L, R, T = A
    n, b, a, dev = bta_shapes(D, L, R, T)
    X2 = _tensor(X, "X").reshape(X.shape[0], -1)
    B2 = _tensor(B, "B").reshape(B.shape[0], -1)
    if X2.shape != B2.shape or X2.shape[0] != n * b + a:
        raise ValueError("X and B must have shape (n*b + a, r)")
    if out is None:
        out = torch.empty_like(B2)
    r = X2.shape[1]
    ws, nbytes = workspace(_lib.OP_RESIDUAL, n, b, a, r, dev)
    rc = _lib.load().rocbta_dresidual(n, b, a, ptr(D), ptr(L), ptr(R), ptr(T), X2.data_ptr(),
                                      B2.data_ptr(), out.data_ptr(), r, r, ws.data_ptr(), nbytes,
                                      stream_of(dev))
    _lib.check(rc, "residual")
    return out


def pobtas_refine(A, F, B, max_iter=10, tol=1e-13):
    """Solve A x = B in place on B with the (mixed) factor F = (D, L, R, T)
    of A and FP64 iterative refinement on the original blocks A.

    Returns the list of relative residual norms ||B - A x||_F / ||B||_F after
    each correction (the first entry is the plain solve).  Stops when it drops
    below tol, after max_iter corrections, or when a correction no longer
    halves it (the FP64 rounding floor of the residual has been reached).
    """
    D, L, R, T = F
    B2 = _tensor(B, "B").reshape(B.shape[0], -1)
    b0 = B2.clone()
    bnorm = float(b0.norm()) or 1.0
    x = b0.clone()
    pobtas(D, L, R, T, x, trans="N")
    pobtas(D, L, R, T, x, trans="C")
    hist = []
    res = torch.empty_like(b0)
    for _ in range(max_iter + 1):
        residual(A, x, b0, out=res)
        rel = float(res.norm()) / bnorm
        if hist and rel > 0.5 * hist[-1]:
            hist.append(rel)
            break
        hist.append(rel)
        if rel < tol or len(hist) > max_iter:
            break
        pobtas(D, L, R, T, res, trans="N")
        pobtas(D, L, R, T, res, trans="C")
        x += res
    B2.copy_(x)
    return hist
