"""Triangular solve with the BTA block Cholesky factor (drop-in for serinv.algs.pobtas).

Reference: src/serinv/algs/pobtas.py:10-62 (signature, kwargs, dispatch),
:65-161 (algorithm) and :164-218 (permuted algorithm, with buffer=); runs as
rocbta_dpobtas / rocbta_dpobtas_permuted.  Like the reference, pobtas does
not validate its keyword arguments.
"""

import torch

from .. import _lib
from ._common import (_tensor, bta_shapes, config, note_status, permutation_buffer, ptr, stream_of,
                      workspace)


def pobtas(L_diagonal_blocks, L_lower_diagonal_blocks, L_lower_arrow_blocks, L_arrow_tip_block, B,
           trans="N", **kwargs):
    """Solve with the block Cholesky factor, in place on B (shape (n*b + a, r) or (n*b + a,)).

    trans "N": B <- L^{-1} B (forward); "T" or "C": B <- L^{-T} B (backward).
    Keyword arguments: partial (default False) skips the last-block and tip
    solves; buffer (from pobtaf(..., buffer=)) selects the permuted variant,
    whose forward pass eliminates the middle blocks 1 .. n-2 into the rows of
    block 0, block n-1 and the tip, and whose backward pass takes the reduced
    system's solution from those rows; device_streaming raises
    NotImplementedError as in the reference (pobtas.py:33-36, 49-52).
    """
    D, L, R, T = L_diagonal_blocks, L_lower_diagonal_blocks, L_lower_arrow_blocks, L_arrow_tip_block
    n, b, a, dev = bta_shapes(D, L, R, T)
    _solve(D, L, R, T, B, trans, kwargs, n, b, a, dev, "pobtas")


def _solve(D, L, R, T, B, trans, kwargs, n, b, a, dev, what):
    partial = bool(kwargs.get("partial", False))
    streaming = kwargs.get("device_streaming", False)
    buf = None
    if kwargs.get("buffer", None) is not None:
        if streaming:  # pobtas.py:31-36, pobts.py:29-32
            raise NotImplementedError("Streaming is not implemented for the permuted arrowhead."
                                      if what == "pobtas" else
                                      "Permuted arrowhead is not implemented.")
        buf = permutation_buffer(kwargs, D)
    elif streaming:
        raise NotImplementedError("Streaming is not implemented for the natural arrowhead.")
    if trans not in ("N", "T", "C"):
        raise ValueError(f"Invalid transpose argument: {trans}.")
    _tensor(B, "B")
    N = n * b + a
    if B.dim() == 1:
        r, ldb = 1, 1
    elif B.dim() == 2:
        r, ldb = B.shape[1], B.shape[1]
    else:
        raise ValueError(f"B must have shape ({N}, r) or ({N},), got {tuple(B.shape)}")
    if B.shape[0] != N:
        raise ValueError(f"B must have {N} rows, got {B.shape[0]}")
    if B.device != dev:
        raise ValueError("B must be on the same device as the factor")
    lib = _lib.load()
    if buf is not None:  # pobtas.py:164-218 / pobts.py:121-165 (partial does not apply)
        ws, nbytes = workspace(_lib.OP_POBTAS_PERMUTED, n, b, a, r, dev)
        rc = lib.rocbta_dpobtas_permuted(n, b, a, ptr(D), ptr(L), ptr(R), ptr(T), ptr(buf), ptr(B),
                                         r, ldb, trans.encode(), ws.data_ptr(), nbytes,
                                         stream_of(dev))
        _lib.check(rc, what)
        return
    ws, nbytes = workspace(_lib.OP_POBTAS, n, b, a, r, dev)
    rc = lib.rocbta_dpobtas(n, b, a, ptr(D), ptr(L), ptr(R), ptr(T), ptr(B), r, ldb,
                            trans.encode(), int(partial), ws.data_ptr(), nbytes, stream_of(dev))
    _lib.check(rc, what)
    # the solve leaves its status in the first int32 of the workspace (0, or
    # -1 after an internal scheduling time-out)
    status = ws[:4].view(torch.int32)
    if not config["check_info"]:
        note_status(status, what)
    elif int(status.item()) < 0:
        raise _lib.RocbtaError(f"{what}: internal scheduling time-out in the solve kernel")
