"""Selected inversion of an SPD BT matrix (drop-in for serinv.algs.pobtsi).

Reference: src/serinv/algs/pobtsi.py:10-75 (signature, kwargs, dispatch),
:78-125 (algorithm) and :128-179 (permuted algorithm, with buffer=); runs as
rocbta_dpobtasi / rocbta_dpobtasi_permuted with a = 0.
"""

from ._common import bta_shapes, check_kwargs, permutation_buffer, streaming_requested
from ._streaming import host_shapes, selinv_streaming
from .pobtasi import _selinv, _selinv_permuted


def pobtsi(L_diagonal_blocks, L_lower_diagonal_blocks, **kwargs):
    """Selected inversion of a block tridiagonal matrix from its factor, in place."""
    check_kwargs(kwargs, {"device_streaming", "buffer", "invert_last_block"})
    invert_last_block = bool(kwargs.get("invert_last_block", True))
    D, L = L_diagonal_blocks, L_lower_diagonal_blocks
    if streaming_requested(kwargs, "pobtsi"):  # pobtsi.py:182-330
        n, b, _ = host_shapes(D, L, None, None)
        selinv_streaming(D, L, None, None, n, b, 0, invert_last_block, "pobtsi")
        return
    n, b, a, dev = bta_shapes(D, L, None, None)
    buf = permutation_buffer(kwargs, D, min_blocks=2)
    if buf is not None:  # pobtsi.py:48-62, 128-179
        _selinv_permuted(D, L, None, None, buf, n, b, 0, dev, "pobtsi")
        return
    _selinv(D, L, None, None, n, b, 0, dev, invert_last_block, "pobtsi")
