"""Non-SPD Schur-complement sweep and selected inversion (drop-ins for
serinv.algs.ddbtasc / ddbtasci / ddbtsc / ddbtsci).

Reference: src/serinv/algs/ddbtasc.py:9-185 (signature, kwargs, dispatch)
and :187-241 (the "AX=I" sweep); ddbtasci.py:9-185 and :187-256;
ddbtsc.py:9-221 and ddbtsci.py:9-237 (the arrowless forms, downward and
upward).  They run as rocbta_dddbtasc / rocbta_dddbtasci: blocked
Gauss-Jordan inverses of the diagonally dominant diagonal blocks (no
pivoting) and grouped GEMMs.  rhs= with quadratic=True ("AXA^T=B",
ddbtasc.py:335-475, ddbtasci.py:343-576; downward) runs as
rocbta_dddbtasc_quadratic / rocbta_dddbtasci_quadratic; the upward BT
quadratic forms (ddbtsc.py:314-356, ddbtsci.py:372-445) run through the
same entry points with direction=1.  rhs= without quadratic (the
reference's "AX=B" branch is an empty placeholder, ddbtasc.py:183-185)
raises NotImplementedError.  buffers= without rhs (ddbtasc.py:244-332,
ddbtasci.py:259-341, ddbtsc.py:224-269, ddbtsci.py:240-297) runs as
rocbta_dddbtasc_permuted / rocbta_dddbtasci_permuted, and with rhs= and
quadratic=True (ddbtasc.py:478-710, ddbtasci.py:578-845, ddbtsc.py:357-466,
ddbtsci.py:446-639) as rocbta_dddbtasc_quadratic_permuted /
rocbta_dddbtasci_quadratic_permuted.
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


_B_KEYS = ("B_diagonal_blocks", "B_lower_diagonal_blocks", "B_upper_diagonal_blocks",
           "B_lower_arrow_blocks", "B_upper_arrow_blocks", "B_arrow_tip_block")


def _rhs_arrays(kwargs, what, arrow):
    """The B blocks of the A X A^T = B form, or None for A X = I, and the
    permutation buffers (Lbuf, Ubuf) or None (ddbtasc.py:74-126 and
    ddbtasci.py:76-110 for the checks and their messages)."""
    rhs = kwargs.get("rhs", None)
    buffers = kwargs.get("buffers", None)
    arrs = None
    if rhs is not None:
        keys = _B_KEYS if arrow else _B_KEYS[:3]
        arrs = [rhs.get(k, None) for k in keys]
        if any(x is None for x in arrs):
            raise ValueError("rhs does not contain the correct arrays")
        if not kwargs.get("quadratic", False):
            raise NotImplementedError(f"{what}: the AX=B form (rhs= without quadratic=True) is not "
                                      "provided")
    bufs = None
    if buffers is not None:
        bufs = [buffers.get(k, None) for k in ("A_lower_buffer_blocks", "A_upper_buffer_blocks")]
        if any(x is None for x in bufs):
            raise ValueError("buffers does not contain the correct arrays for A")
        if arrs is not None:
            bb = [buffers.get(k, None) for k in ("B_lower_buffer_blocks", "B_upper_buffer_blocks")]
            if any(x is None for x in bb):
                raise ValueError("buffers does not contain the correct arrays for B")
            bufs += bb
    return arrs, bufs


def _check_like(B, A, name):
    _tensor(B, name)
    if tuple(B.shape) != tuple(A.shape) or B.device != A.device:
        raise ValueError(f"{name} must match the shape {tuple(A.shape)} and device of A's block")


def _direction(kwargs):
    d = kwargs.get("direction", "downward")
    if d == "downward":
        return 0
    if d == "upward":
        return 1
    # ddbtsc.py:86-89
    raise ValueError("Optional keyword argument `direction` must be either 'downward' or 'upward'")


def _run(fn_name, op, D, Lo, Up, Rl, Ru, T, n, b, a, dev, invert_last_block, direction, what,
         B=None, bufs=None):
    lib = _lib.load()
    if bufs is not None:  # permuted: the middle blocks 1 .. n-2
        if n < 2:
            raise IndexError(f"{what}: the permuted form needs at least 2 diagonal blocks, got {n}")
        for x, nm in zip(bufs, ("A_lower_buffer_blocks", "A_upper_buffer_blocks",
                                "B_lower_buffer_blocks", "B_upper_buffer_blocks")):
            _check_like(x, Lo, nm)
        ws, nbytes = workspace(op, n, b, a, 0, dev)
        if B is None:
            rc = getattr(lib, fn_name + "_permuted")(
                n, b, a, ptr(D), ptr(Lo), ptr(Up), ptr(Rl), ptr(Ru), ptr(T), ptr(bufs[0]),
                ptr(bufs[1]), ws.data_ptr(), nbytes, stream_of(dev))
        else:
            for x, y, nm in zip(B, (D, Lo, Up, Rl, Ru, T), _B_KEYS):
                _check_like(x, y, nm)
            Bp = [ptr(x) for x in B] + [None] * (6 - len(B))
            rc = getattr(lib, fn_name + "_quadratic_permuted")(
                n, b, a, ptr(D), ptr(Lo), ptr(Up), ptr(Rl), ptr(Ru), ptr(T), ptr(bufs[0]),
                ptr(bufs[1]), *Bp, ptr(bufs[2]), ptr(bufs[3]), ws.data_ptr(), nbytes,
                stream_of(dev))
        _lib.check(rc, what)
        return
    ws, nbytes = workspace(op, n, b, a, 0, dev)
    if B is not None:  # A X A^T = B
        for x, y, nm in zip(B, (D, Lo, Up, Rl, Ru, T), _B_KEYS):
            if y is not None:
                _check_like(x, y, nm)
        Bp = [ptr(x) for x in B] + [None] * (6 - len(B))
        rc = getattr(lib, fn_name + "_quadratic")(
            n, b, a, ptr(D), ptr(Lo), ptr(Up), ptr(Rl), ptr(Ru), ptr(T), *Bp,
            int(invert_last_block), int(direction), ws.data_ptr(), nbytes, stream_of(dev))
        _lib.check(rc, what)
        return
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
    B, bufs = _rhs_arrays(kwargs, "ddbtasc", True)
    arrs = (A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks,
            A_lower_arrow_blocks, A_upper_arrow_blocks, A_arrow_tip_block)
    n, b, a, dev = _shapes(*arrs)
    _run("rocbta_dddbtasc", _lib.OP_DDBTASC, *arrs, n, b, a, dev,
         bool(kwargs.get("invert_last_block", True)), 0, "ddbtasc", B, bufs)


def ddbtasci(A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks,
             A_lower_arrow_blocks, A_upper_arrow_blocks, A_arrow_tip_block, **kwargs):
    """Selected inversion from the ddbtasc output, in place: the blocks of
    A^{-1} on the BTA pattern."""
    check_kwargs(kwargs, {"rhs", "quadratic", "buffers", "invert_last_block"})
    B, bufs = _rhs_arrays(kwargs, "ddbtasci", True)
    arrs = (A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks,
            A_lower_arrow_blocks, A_upper_arrow_blocks, A_arrow_tip_block)
    n, b, a, dev = _shapes(*arrs)
    _run("rocbta_dddbtasci", _lib.OP_DDBTASCI, *arrs, n, b, a, dev,
         bool(kwargs.get("invert_last_block", True)), 0, "ddbtasci", B, bufs)


def ddbtsc(A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks, **kwargs):
    """Schur-complement sweep of a diagonally dominant BT matrix, in place;
    direction "downward" (default) or "upward"."""
    check_kwargs(kwargs, {"rhs", "quadratic", "buffers", "invert_last_block", "direction"})
    B, bufs = _rhs_arrays(kwargs, "ddbtsc", False)
    direction = _direction(kwargs)
    D, Lo, Up = A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks
    n, b, a, dev = _shapes(D, Lo, Up, None, None, None)
    _run("rocbta_dddbtasc", _lib.OP_DDBTASC, D, Lo, Up, None, None, None, n, b, 0, dev,
         bool(kwargs.get("invert_last_block", True)), direction, "ddbtsc", B, bufs)


def ddbtsci(A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks, **kwargs):
    """Selected inversion from the ddbtsc output (same direction), in place."""
    check_kwargs(kwargs, {"rhs", "quadratic", "buffers", "direction"})
    B, bufs = _rhs_arrays(kwargs, "ddbtsci", False)
    direction = _direction(kwargs)
    D, Lo, Up = A_diagonal_blocks, A_lower_diagonal_blocks, A_upper_diagonal_blocks
    n, b, a, dev = _shapes(D, Lo, Up, None, None, None)
    _run("rocbta_dddbtasci", _lib.OP_DDBTASCI, D, Lo, Up, None, None, None, n, b, 0, dev, True,
         direction, "ddbtsci", B, bufs)
