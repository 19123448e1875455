"""Selected inversion of an SPD BT matrix (drop-in for serinv.algs.pobtsi).

Reference: src/serinv/algs/pobtsi.py:10-75 (signature, kwargs, dispatch) and
:78-125 (algorithm); runs as rocbta_dpobtasi with a = 0.
"""

from ._common import bta_shapes, check_kwargs, reject_unsupported
from .pobtasi import _selinv


def pobtsi(L_diagonal_blocks, L_lower_diagonal_blocks, **kwargs):
    """Selected inversion of a block tridiagonal matrix from its factor, in place."""
    check_kwargs(kwargs, {"device_streaming", "buffer", "invert_last_block"})
    reject_unsupported(kwargs, "pobtsi")
    invert_last_block = bool(kwargs.get("invert_last_block", True))
    D, L = L_diagonal_blocks, L_lower_diagonal_blocks
    n, b, a, dev = bta_shapes(D, L, None, None)
    _selinv(D, L, None, None, n, b, 0, dev, invert_last_block, "pobtsi")
