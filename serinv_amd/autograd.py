"""Differentiable log-determinant and quadratic form of an SPD BTA matrix.

No counterpart in the reference (an extension for Laplace-approximation and
Gaussian-likelihood outer loops).  For A given by its stored blocks
(D, L, R, T) in serinv's layout and right-hand sides y (N, r) or (N,):

    logdet, quad = pobta_logdet_quad(D, L, R, T, y)
    logdet = log det A,   quad = sum_k y_k^T A^{-1} y_k

Forward: pobtaf on copies of the blocks, rocbta_dlogdet, two pobtas sweeps
(x = A^{-1} y) and rocbta_dquad.  Backward: pobtasi on the saved factor gives
the selected inverse X = A^{-1} on the BTA pattern, and the gradient kernel
rocbta_dgrad forms

    dlogdet/dA = X,   dquad/dA = -x x^T,   dquad/dy = 2 x

mapped to the stored entries (only the lower triangles of D and T are read
by the factorization, so their strict-lower entries carry weight 2 and the
upper ones 0; L and R blocks carry weight 2).  With symmetric D and T blocks
that depend on parameters theta, autograd through them yields, for
f = logdet + quad, df/dtheta = tr(A^{-1} dA/dtheta) - x^T (dA/dtheta) x.
"""

import torch

from . import _lib
from .algs import pobtaf, pobtas, pobtasi
from .algs._common import _tensor, bta_shapes, ptr, stream_of
from .logdet import pobta_logdet


def _quad(y, x):
    out = torch.empty((), dtype=torch.float64, device=y.device)
    rows, r = y.shape
    rc = _lib.load().rocbta_dquad(rows, r, y.data_ptr(), x.data_ptr(), r, out.data_ptr(),
                                  stream_of(y.device))
    _lib.check(rc, "pobta_logdet_quad")
    return out


class _LogdetQuad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D, L, R, T, y):
        Df, Lf, Rf, Tf = (t.detach().clone() for t in (D, L, R, T))
        pobtaf(Df, Lf, Rf, Tf)
        logdet = pobta_logdet(Df, Tf)
        y2 = y.detach().reshape(y.shape[0], -1).contiguous()
        x = y2.clone()
        pobtas(Df, Lf, Rf, Tf, x, trans="N")
        pobtas(Df, Lf, Rf, Tf, x, trans="C")
        quad = _quad(y2, x)
        ctx.save_for_backward(Df, Lf, Rf, Tf, x)
        ctx.y_shape = y.shape
        return logdet, quad

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, g_logdet, g_quad):
        if getattr(ctx, "consumed", False):
            raise RuntimeError("pobta_logdet_quad: backward runs the selected inversion in place on "
                               "the saved factor, so it can run only once per forward")
        ctx.consumed = True
        Df, Lf, Rf, Tf, x = ctx.saved_tensors
        n, b, a = Df.shape[0], Df.shape[1], Rf.shape[1]
        dev = Df.device
        # selected inverse in place on the saved factor (used once)
        pobtasi(Df, Lf, Rf, Tf)
        gl = None if g_logdet is None else g_logdet.detach().to(torch.float64).reshape(()).contiguous()
        gq = None if g_quad is None else g_quad.detach().to(torch.float64).reshape(()).contiguous()
        need_y = ctx.needs_input_grad[4] and gq is not None
        gy = torch.empty_like(x) if need_y else None
        r = x.shape[1]
        rc = _lib.load().rocbta_dgrad(
            n, b, a, ptr(Df), ptr(Lf), ptr(Rf), ptr(Tf), ptr(x), r, r,
            None if gl is None else gl.data_ptr(), None if gq is None else gq.data_ptr(),
            ptr(Df), ptr(Lf), ptr(Rf), ptr(Tf), None if gy is None else gy.data_ptr(),
            stream_of(dev))
        _lib.check(rc, "pobta_logdet_quad backward")
        gyo = gy.reshape(ctx.y_shape) if gy is not None else None
        return Df, Lf, Rf, Tf, gyo


def pobta_logdet_quad(D, L, R, T, y):
    """Return (log det A, sum_k y_k^T A^{-1} y_k) as 0-dim float64 tensors,
    differentiable with respect to D, L, R, T and y (float64 ROCm tensors in
    serinv's block layout; R of shape (n, a, b), T of shape (a, a), a >= 0)."""
    n, b, a, dev = bta_shapes(D, L, R, T)
    _tensor(y, "y")
    if y.shape[0] != n * b + a or y.dim() not in (1, 2):
        raise ValueError(f"y must have shape ({n * b + a},) or ({n * b + a}, r)")
    return _LogdetQuad.apply(D, L, R, T, y)
