"""Block products on the BTA pattern, in FP64 torch on the GPU.

Size-independent property checks of the three routines at full size, where
no dense matrix fits:

* llt_pattern(F): the blocks of L L^T on the BTA pattern, from the factor
  F = (L_D, L_L, L_R, L_T) written by pobtaf.  They must reproduce A.
* ax_diag(A, X): the diagonal blocks and the tip block of A X, from the
  original blocks A = (D, L, R, T) and the selected inverse X written by
  pobtasi.  Only blocks of X on the pattern enter them:
    (AX)_ii = A_{i,i-1} X_{i,i-1}^T + A_ii X_ii + A_{i+1,i}^T X_{i+1,i}
              + A_{t,i}^T X_{t,i}
    (AX)_tt = sum_i A_{t,i} X_{t,i}^T + A_tt X_tt
  They must be identity blocks.
* bta_matvec(A, x): A x from the blocks (D and T hold full symmetric blocks).

Transposed operands are made contiguous before every batched product.
"""

import torch


def _t(x):
    return x.transpose(-1, -2).contiguous()


def llt_pattern(F):
    D, L, R, T = F
    n, b = D.shape[0], D.shape[1]
    a = R.shape[1]
    PD = torch.bmm(D, _t(D))
    if n > 1:
        PD[1:] += torch.bmm(L, _t(L))
        PL = torch.bmm(L, _t(D[:-1]))
    else:
        PL = torch.empty_like(L)
    PR = torch.bmm(R, _t(D))
    if n > 1:
        PR[1:] += torch.bmm(R[:-1], _t(L))
    Rf = R.permute(1, 0, 2).reshape(a, n * b)
    PT = Rf @ Rf.T + T @ T.T
    return PD, PL, PR, PT


def ax_diag(A, X):
    D0, L0, R0, T0 = A
    XD, XL, XR, XT = X
    n = D0.shape[0]
    a = R0.shape[1]
    P = torch.bmm(D0, XD)
    if n > 1:
        P[1:] += torch.bmm(L0, _t(XL))
        P[:-1] += torch.bmm(_t(L0), XL)
    if a:
        P += torch.bmm(_t(R0), XR)
        Rf = R0.permute(1, 0, 2).reshape(a, -1)
        Xf = XR.permute(1, 0, 2).reshape(a, -1)
        PT = Rf @ Xf.T + T0 @ XT
    else:
        PT = torch.zeros((0, 0), dtype=D0.dtype, device=D0.device)
    return P, PT


def bta_matvec(A, x):
    D0, L0, R0, T0 = A
    n, b = D0.shape[0], D0.shape[1]
    a = R0.shape[1]
    x2 = x.reshape(x.shape[0], -1)
    y = torch.zeros_like(x2)
    xb = x2[: n * b].view(n, b, -1)
    yb = y[: n * b].view(n, b, -1)
    yb.copy_(torch.bmm(D0, xb))
    if n > 1:
        yb[1:] += torch.bmm(L0, xb[:-1])
        yb[:-1] += torch.bmm(_t(L0), xb[1:])
    if a:
        xt = x2[n * b:]
        Rf = R0.permute(1, 0, 2).reshape(a, n * b)
        y[n * b:] = Rf @ x2[: n * b] + T0 @ xt
        y[: n * b] += (Rf.T @ xt)
    return y.reshape(x.shape)


def maxrel(x, ref):
    """max |x - ref| / max |ref| (the parity measure; 0 for empty arrays)."""
    if ref.numel() == 0:
        return 0.0
    return float((x - ref).abs().max() / ref.abs().max())


def identity_err(P):
    """max |P_k - I| over a stack of square blocks (or one block)."""
    if P.numel() == 0:
        return 0.0
    m = P.shape[-1]
    eye = torch.eye(m, dtype=P.dtype, device=P.device)
    return float((P - eye).abs().max())
