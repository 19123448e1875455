"""Block Cholesky factorization of an SPD BT matrix (drop-in for serinv.algs.pobtf).

Reference: src/serinv/algs/pobtf.py:11-80 (signature, kwargs, dispatch) and
:83-122 (algorithm); runs as rocbta_dpobtaf with a = 0.
"""

from ._common import bta_shapes, check_kwargs, reject_unsupported
from .pobtaf import _factor


def pobtf(A_diagonal_blocks, A_lower_diagonal_blocks, **kwargs):
    """Cholesky factorization of a block tridiagonal matrix, in place.

    diagonal blocks (n, b, b) -> L_ii (upper triangle zero); lower diagonal
    blocks (n-1, b, b) -> L_{i+1,i}.  Keyword arguments as serinv's pobtf.
    """
    check_kwargs(kwargs, {"device_streaming", "buffer", "factorize_last_block"})
    reject_unsupported(kwargs, "pobtf")
    factorize_last_block = bool(kwargs.get("factorize_last_block", True))
    D, L = A_diagonal_blocks, A_lower_diagonal_blocks
    n, b, a, dev = bta_shapes(D, L, None, None)
    _factor(D, L, None, None, n, b, 0, dev, factorize_last_block, "pobtf")
