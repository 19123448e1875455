"""log-determinant of an SPD BTA / BT matrix from its block Cholesky factor.

No counterpart in the reference (an extension for likelihood evaluation):
    log det A = 2 * sum_i sum_j log (L_ii)_jj + 2 * sum_j log (L_T)_jj
evaluated on the GPU by rocbta_dlogdet from the arrays written by pobtaf /
pobtf.  Returns a 0-dim float64 tensor on the factor's device.
"""

import torch

from . import _lib
from .algs._common import _tensor, stream_of, workspace


def pobta_logdet(L_diagonal_blocks, L_arrow_tip_block=None):
    D = _tensor(L_diagonal_blocks, "diagonal blocks")
    if D.dim() != 3 or D.shape[1] != D.shape[2]:
        raise ValueError("diagonal blocks must have shape (n, b, b)")
    n, b = D.shape[0], D.shape[1]
    a = 0
    T = L_arrow_tip_block
    if T is not None:
        _tensor(T, "arrow tip block")
        a = T.shape[0]
    out = torch.empty((), dtype=torch.float64, device=D.device)
    ws, nbytes = workspace(_lib.OP_LOGDET, n, b, a, 0, D.device)
    rc = _lib.load().rocbta_dlogdet(n, b, a, D.data_ptr(), T.data_ptr() if a else None,
                                    out.data_ptr(), ws.data_ptr(), nbytes, stream_of(D.device))
    _lib.check(rc, "pobta_logdet")
    return out
