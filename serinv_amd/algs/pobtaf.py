"""Block Cholesky factorization of an SPD BTA matrix (drop-in for serinv.algs.pobtaf).

Reference: src/serinv/algs/pobtaf.py:11-94 (signature, kwargs, dispatch),
:97-183 (natural-ordering algorithm) and :186-281 (permuted algorithm, with
buffer=).  The factorization runs on the GPU as hand-written gfx950 HIP
kernels (librocbta.so, rocbta_dpobtaf / rocbta_dpobtaf_permuted).
"""

import torch

from .. import _lib
from ._common import (bta_shapes, check_kwargs, config, note_status, permutation_buffer, ptr,
                      raise_on_info, stream_of, streaming_requested, workspace)
from ._streaming import factor_streaming, host_shapes


def pobtaf(A_diagonal_blocks, A_lower_diagonal_blocks, A_lower_arrow_blocks, A_arrow_tip_block,
           **kwargs):
    """Cholesky factorization of a block tridiagonal with arrowhead matrix, in place.

    Parameters follow serinv: diagonal blocks (n, b, b), lower diagonal blocks
    (n-1, b, b), lower arrow blocks (n, a, b), arrow tip (a, a); float64 torch
    tensors on a ROCm device.  On exit they hold L_ii (upper triangle zero),
    L_{i+1,i}, L_{t,i} and L_T.

    Keyword arguments: factorize_last_block (default True); buffer (a
    tensor like A_diagonal_blocks, see
    serinv_amd.utils.allocate_pobtax_permutation_buffers) selects the
    permuted variant, which factors the middle blocks 1 .. n-2 and leaves the
    reduced system (blocks 0 and n-1, buffer[n-1], the tip) to the caller;
    device_streaming=True takes HOST tensors (float64, CPU, best pinned) and
    streams the blocks through the GPU (pobtaf.py:284-524).
    A non-positive pivot raises numpy.linalg.LinAlgError (when
    serinv_amd.algs.config["check_info"] is True).
    """
    check_kwargs(kwargs, {"device_streaming", "buffer", "factorize_last_block"})
    factorize_last_block = bool(kwargs.get("factorize_last_block", True))
    D, L, R, T = A_diagonal_blocks, A_lower_diagonal_blocks, A_lower_arrow_blocks, A_arrow_tip_block
    if streaming_requested(kwargs, "pobtaf"):
        n, b, a = host_shapes(D, L, R, T)
        factor_streaming(D, L, R, T, n, b, a, factorize_last_block, "pobtaf")
        return
    n, b, a, dev = bta_shapes(D, L, R, T)
    buf = permutation_buffer(kwargs, D, min_blocks=2)
    if buf is not None:
        _factor_permuted(D, L, R, T, buf, n, b, a, dev, "pobtaf")
        return
    _factor(D, L, R, T, n, b, a, dev, factorize_last_block, "pobtaf")


def _factor(D, L, R, T, n, b, a, dev, factorize_last_block, what, mixed=False):
    lib = _lib.load()
    ws, nbytes = workspace(_lib.OP_POBTAF, n, b, a, 0, dev)
    info = torch.empty(1, dtype=torch.int32, device=dev)
    fn = lib.rocbta_mpobtaf if mixed else lib.rocbta_dpobtaf
    rc = fn(n, b, a, ptr(D), ptr(L), ptr(R), ptr(T), int(factorize_last_block),
            info.data_ptr(), ws.data_ptr(), nbytes, stream_of(dev))
    _lib.check(rc, what)
    if config["check_info"]:
        raise_on_info(info, what)
    else:
        note_status(info, what)


def _factor_permuted(D, L, R, T, buf, n, b, a, dev, what):
    lib = _lib.load()
    ws, nbytes = workspace(_lib.OP_POBTAF_PERMUTED, n, b, a, 0, dev)
    info = torch.empty(1, dtype=torch.int32, device=dev)
    rc = lib.rocbta_dpobtaf_permuted(n, b, a, ptr(D), ptr(L), ptr(R), ptr(T), ptr(buf),
                                     info.data_ptr(), ws.data_ptr(), nbytes, stream_of(dev))
    _lib.check(rc, what)
    if config["check_info"]:
        raise_on_info(info, what)
    else:
        note_status(info, what)
