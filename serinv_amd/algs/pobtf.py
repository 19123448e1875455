"""Block Cholesky factorization of an SPD BT matrix (drop-in for serinv.algs.pobtf).

Reference: src/serinv/algs/pobtf.py:11-80 (signature, kwargs, dispatch),
:83-122 (algorithm) and :125-185 (permuted algorithm, with buffer=); runs as
rocbta_dpobtaf / rocbta_dpobtaf_permuted with a = 0.
"""

from ._common import bta_shapes, check_kwargs, permutation_buffer, streaming_requested
from ._streaming import factor_streaming, host_shapes
from .pobtaf import _factor, _factor_permuted


def pobtf(A_diagonal_blocks, A_lower_diagonal_blocks, **kwargs):
    """Cholesky factorization of a block tridiagonal matrix, in place.

    diagonal blocks (n, b, b) -> L_ii (upper triangle zero); lower diagonal
    blocks (n-1, b, b) -> L_{i+1,i}.  Keyword arguments as serinv's pobtf.
    """
    check_kwargs(kwargs, {"device_streaming", "buffer", "factorize_last_block"})
    factorize_last_block = bool(kwargs.get("factorize_last_block", True))
    D, L = A_diagonal_blocks, A_lower_diagonal_blocks
    if streaming_requested(kwargs, "pobtf"):  # pobtf.py:188-329
        n, b, _ = host_shapes(D, L, None, None)
        factor_streaming(D, L, None, None, n, b, 0, factorize_last_block, "pobtf")
        return
    n, b, a, dev = bta_shapes(D, L, None, None)
    buf = permutation_buffer(kwargs, D, min_blocks=2)
    if buf is not None:  # pobtf.py:53-68, 125-185
        _factor_permuted(D, L, None, None, buf, n, b, 0, dev, "pobtf")
        return
    _factor(D, L, None, None, n, b, 0, dev, factorize_last_block, "pobtf")
