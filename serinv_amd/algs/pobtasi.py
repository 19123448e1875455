"""Selected inversion of an SPD BTA matrix (drop-in for serinv.algs.pobtasi).

Reference: src/serinv/algs/pobtasi.py:10-89 (signature, kwargs, dispatch),
:92-172 (natural-ordering algorithm) and :175-244 (permuted algorithm, with
buffer=); runs as rocbta_dpobtasi / rocbta_dpobtasi_permuted.
"""

from .. import _lib
from ._common import (bta_shapes, check_kwargs, permutation_buffer, ptr, stream_of,
                      streaming_requested, workspace)
from ._streaming import host_shapes, selinv_streaming


def pobtasi(L_diagonal_blocks, L_lower_diagonal_blocks, L_lower_arrow_blocks, L_arrow_tip_block,
            **kwargs):
    """Selected inversion from the block Cholesky factor, in place.

    On exit the arrays hold X_ii (full symmetric), X_{i+1,i}, X_{t,i} and X_T,
    the blocks of A^{-1} on the BTA pattern.  Keyword arguments:
    invert_last_block (default True); buffer selects the permuted variant
    (the factor from pobtaf(..., buffer=) with the reduced system's blocks
    already inverted; on exit buffer[i] = X_{0,i}); device_streaming=True
    takes HOST tensors and streams the blocks through the GPU
    (pobtasi.py:247-496).
    """
    check_kwargs(kwargs, {"device_streaming", "buffer", "invert_last_block"})
    invert_last_block = bool(kwargs.get("invert_last_block", True))
    D, L, R, T = L_diagonal_blocks, L_lower_diagonal_blocks, L_lower_arrow_blocks, L_arrow_tip_block
    if streaming_requested(kwargs, "pobtasi"):
        n, b, a = host_shapes(D, L, R, T)
        selinv_streaming(D, L, R, T, n, b, a, invert_last_block, "pobtasi")
        return
    n, b, a, dev = bta_shapes(D, L, R, T)
    buf = permutation_buffer(kwargs, D, min_blocks=2)
    if buf is not None:
        _selinv_permuted(D, L, R, T, buf, n, b, a, dev, "pobtasi")
        return
    _selinv(D, L, R, T, n, b, a, dev, invert_last_block, "pobtasi")


def _selinv(D, L, R, T, n, b, a, dev, invert_last_block, what, mixed=False):
    lib = _lib.load()
    ws, nbytes = workspace(_lib.OP_POBTASI, n, b, a, 0, dev)
    fn = lib.rocbta_mpobtasi if mixed else lib.rocbta_dpobtasi
    rc = fn(n, b, a, ptr(D), ptr(L), ptr(R), ptr(T), int(invert_last_block),
            ws.data_ptr(), nbytes, stream_of(dev))
    _lib.check(rc, what)


def _selinv_permuted(D, L, R, T, buf, n, b, a, dev, what):
    lib = _lib.load()
    ws, nbytes = workspace(_lib.OP_POBTASI_PERMUTED, n, b, a, 0, dev)
    rc = lib.rocbta_dpobtasi_permuted(n, b, a, ptr(D), ptr(L), ptr(R), ptr(T), ptr(buf),
                                      ws.data_ptr(), nbytes, stream_of(dev))
    _lib.check(rc, what)
