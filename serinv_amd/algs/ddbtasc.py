"""Non-SPD Schur-complement sweep and selected inversion (drop-ins for
serinv.algs.ddbtasc / ddbtasci / ddbtsc / ddbtsci).

Reference: src/serinv/algs/ddbtasc.py:9-185 (signature, kwargs, dispatch)
and :187-241 (the "AX=I" sweep); ddbtasci.py:9-185 and :187-256;
ddbtsc.py:9-221 and ddbtsci.py:9-237 (the arrowless forms, downward and
upward).  They run as rocbta_dddbtasc / rocbta_dddbtasci: blocked
Gauss-Jordan inverses of the diagonally dominant diagonal blocks (no
pivoting) and grouped GEMMs.  The rhs= / quadratic= ("AXA^T=B") and
buffers= (permuted) forms raise NotImplementedError.
"""

import torch

from .. import _lib
from ._common import _tensor, check_kwargs, ptr, stream_of, workspace


def _shapes(D, Lo, Up, Rl, Ru, T):
    _tensor(D, "diagonal blocks")
    if D.dim() != 3 or D.shape[1] != D.shape[2]:
        raise ValueError(f"diagonal blocks must have shape (n, b, b), got {tuple(D.shape)}")
    n, b = D.shape[0], D.shape[1]
    if n < 1 or b < 1:
        raise ValueError("need at least one diagonal block of size >= 1")
    for x, name in ((Lo, "lower diagonal blocks"), (Up, "upper diagonal blocks")):
        _tensor(x, name)
        if tuple(x.shape) != (n - 1, b, b):
            raise ValueError(f"{name} must have shape {(n - 1, b, b)}, got {tuple(x.shape)}")
    a = 0
    if Rl is not None:
        _tensor(Rl, "lower arrow blocks")
        if Rl.dim() != 3 or Rl.shape[0] != n or Rl.shape[2] != b:
            raise ValueError(f"lower arrow blocks must have shape (n, a, {b}), got {tuple(Rl.shape)}")
        a = Rl.shape[1]
        _tensor(Ru, "upper arrow blocks")
        if tuple(Ru.shape) != (n, b, a):
            raise ValueError(f"upper arrow blocks must have shape {(n, b, a)}, got {tuple(Ru.shape)}")
        _tensor(T, "arrow tip block")
        if tuple(T.shape) != (a, a):
            raise ValueError(f"arrow tip block must have shape {(a, a)}, got {tuple(T.shape)}")
    devs = {t.device for t in (D, Lo, Up, Rl, Ru, T) if t is not None}
    if len(devs) != 1:
        raise ValueError(f"all blocks must be on the same device, got {devs}")
    return n, b, a, D.device


def _reject_forms(kwargs, what):
    # ddbtasc.py:74-185: rhs (AX=B / AXA^T=B) and buffers (permuted) forms
    if kwargs.get("rhs", None) is not None:
        raise NotImplementedError(f"{what}: the rhs= (AX=B / AXA^T=B) forms are not provided")
    if kwargs.get("buffers", None) is not None:
        raise NotImplementedError(f"{what}: the permuted (buffers=) form is not provided")


def _direction(kwargs):
    d = kwargs.get("direction", "downward")
    if d == "downward":
        return 0
    if d == "upward":
        return 1
    # ddbtsc.py:86-89
    raise ValueError("Optional keyword argument `direction` must be either 'downward' or 'upward'")


def _run(fn_name, op, D, Lo, Up, Rl, Ru, T, n, b, a, dev, invert_last_block, direction, what):
    lib = _lib.load()
    ws, nbytes = workspace(op, n, b, a, 0, dev)
    rc = getattr(lib, fn_name)(n, b, a, ptr(D), ptr(Lo), ptr(Up), ptr(Rl), ptr(Ru), ptr(T),
                               int(invert_last_block), int(direction), ws.data_ptr(), nbytes,
                               stream_of(dev))
    _lib.check(rc, what)


def ddbtasc(A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks,
            A_lower_arrow_blocks, A_upper_arrow_blocks, A_arrow_tip_block, **kwargs):
    """Schur-complement sweep of a diagonally dominant BTA matrix, in place.

    On exit the diagonal blocks hold the inverses of the Schur complements,
    the arrow blocks and the tip their Schur-updated values (the tip
    inverted with invert_last_block=True, the default)."""
    check_kwargs(kwargs, {"rhs", "quadratic", "buffers", "invert_last_block"})
    _reject_forms(kwargs, "ddbtasc")
    arrs = (A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks,
            A_lower_arrow_blocks, A_upper_arrow_blocks, A_arrow_tip_block)
    n, b, a, dev = _shapes(*arrs)
    _run("rocbta_dddbtasc", _lib.OP_DDBTASC, *arrs, n, b, a, dev,
         bool(kwargs.get("invert_last_block", True)), 0, "ddbtasc")


def ddbtasci(A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks,
             A_lower_arrow_blocks, A_upper_arrow_blocks, A_arrow_tip_block, **kwargs):
    """Selected inversion from the ddbtasc output, in place: the blocks of
    A^{-1} on the BTA pattern."""
    check_kwargs(kwargs, {"rhs", "quadratic", "buffers", "invert_last_block"})
    _reject_forms(kwargs, "ddbtasci")
    arrs = (A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks,
            A_lower_arrow_blocks, A_upper_arrow_blocks, A_arrow_tip_block)
    n, b, a, dev = _shapes(*arrs)
    _run("rocbta_dddbtasci", _lib.OP_DDBTASCI, *arrs, n, b, a, dev,
         bool(kwargs.get("invert_last_block", True)), 0, "ddbtasci")


def ddbtsc(A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks, **kwargs):
    """Schur-complement sweep of a diagonally dominant BT matrix, in place;
    direction "downward" (default) or "upward"."""
    check_kwargs(kwargs, {"rhs", "quadratic", "buffers", "invert_last_block", "direction"})
    _reject_forms(kwargs, "ddbtsc")
    direction = _direction(kwargs)
    D, Lo, Up = A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks
    n, b, a, dev = _shapes(D, Lo, Up, None, None, None)
    _run("rocbta_dddbtasc", _lib.OP_DDBTASC, D, Lo, Up, None, None, None, n, b, 0, dev,
         bool(kwargs.get("invert_last_block", True)), direction, "ddbtsc")


def ddbtsci(A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks, **kwargs):
    """Selected inversion from the ddbtsc output (same direction), in place."""
    check_kwargs(kwargs, {"rhs", "quadratic", "buffers", "direction"})
    _reject_forms(kwargs, "ddbtsci")
    direction = _direction(kwargs)
    D, Lo, Up = A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks
    n, b, a, dev = _shapes(D, Lo, Up, None, None, None)
    _run("rocbta_dddbtasci", _lib.OP_DDBTASCI, D, Lo, Up, None, None, None, n, b, 0, dev, True,
         direction, "ddbtsci")
