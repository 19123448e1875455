"""Shared argument handling for the drop-in routines.

Arrays are torch tensors on a ROCm device (float64, C-contiguous) in serinv's
block layout; the routines overwrite them in place and return None, like the
reference (src/serinv/algs/*.py).  There is no host path and no fallback: a
tensor that is not on the GPU, or a missing librocbta.so, raises.
"""

import threading

import numpy as np
import torch

from .. import _lib

# Runtime switches (module level, so that the reference's call signatures stay
# unchanged).  check_info: after a factorization, read the device pivot status
# back and raise numpy.linalg.LinAlgError on a non-positive pivot, as the
# reference's NumPy path does (scipy.linalg.cholesky raises).  Turning it off
# keeps the call fully asynchronous, as the reference's CuPy path is; the
# status words of such calls (pivot failures and internal scheduling
# time-outs, after which the outputs are partial) stay on the device and
# check_status() reports them later.
config = {"check_info": True}

_pending = []  # (device int32 status tensor, routine name) of unchecked calls
_pending_lock = threading.Lock()
_PENDING_MAX = 256


def note_status(status, what):
    """Keep the device status word of an unchecked call for check_status()."""
    with _pending_lock:
        _pending.append((status.reshape(-1)[:1].clone(), what))
        if len(_pending) > _PENDING_MAX:
            # fold into two entries (smallest and largest status) without a sync
            allv = torch.cat([p[0].to(torch.int64) for p in _pending])
            names = sorted({p[1] for p in _pending})
            _pending[:] = [(allv.min().reshape(1).to(torch.int32), "/".join(names)),
                           (allv.max().reshape(1).to(torch.int32), "/".join(names))]


def check_status():
    """Synchronize on and clear the status words of every call made with
    config["check_info"] = False, and raise like a checked call would on the
    first failure: RocbtaError after an internal scheduling time-out (the
    outputs of that call are then partial), numpy.linalg.LinAlgError after a
    non-positive pivot."""
    with _pending_lock:
        items = list(_pending)
        _pending.clear()
    for st, what in items:
        v = int(st.item())
        if v < 0:
            raise _lib.RocbtaError(f"{what}: internal scheduling time-out (outputs incomplete)")
        if v > 0:
            raise np.linalg.LinAlgError(
                f"{what}: the leading minor of order {v} is not positive definite")


def check_kwargs(kwargs, expected):
    # src/serinv/algs/pobtaf.py:50-53 (same pattern in pobtasi/pobtf/pobtsi)
    unexpected = set(kwargs) - set(expected)
    if unexpected:
        raise TypeError(f"Unexpected keyword arguments: {unexpected}")


def reject_unsupported(kwargs, what):
    """For the extensions without permuted or streaming forms (mixed)."""
    if kwargs.get("buffer", None) is not None:
        raise NotImplementedError(f"{what}: the permuted (buffer=...) form is not provided")
    if kwargs.get("device_streaming", False):
        raise NotImplementedError(f"{what}: host-device streaming is not provided")


def streaming_requested(kwargs, what):
    """device_streaming=True: the host-streaming form (natural ordering).
    The permuted streaming forms (buffer= with device_streaming=True) are
    not provided."""
    if not kwargs.get("device_streaming", False):
        return False
    if kwargs.get("buffer", None) is not None:
        raise NotImplementedError(f"{what}: host-device streaming of the permuted form is not "
                                  "implemented")
    return True


def permutation_buffer(kwargs, D, min_blocks=1):
    """The buffer= argument of the permuted variants, validated: a tensor
    like the diagonal blocks (serinv.utils.allocate_pobtax_permutation_buffers
    returns empty_like(A_diagonal_blocks)), or None for the natural ordering.
    The reference indexes buffer[1] (and L[0]), which fails for fewer than
    min_blocks diagonal blocks with an IndexError; so does this."""
    buf = kwargs.get("buffer", None)
    if buf is None:
        return None
    _tensor(buf, "buffer")
    if tuple(buf.shape) != tuple(D.shape):
        raise ValueError(f"buffer must have the shape of the diagonal blocks {tuple(D.shape)}, "
                         f"got {tuple(buf.shape)}")
    if buf.device != D.device:
        raise ValueError("buffer must be on the same device as the blocks")
    if D.shape[0] < min_blocks:
        raise IndexError(f"the permuted variant needs at least {min_blocks} diagonal blocks, "
                         f"got {D.shape[0]}")
    return buf


def _tensor(x, name):
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor on a ROCm device, got {type(x).__name__}")
    if x.device.type != "cuda":
        raise TypeError(f"{name} must live on the GPU (got device {x.device})")
    if x.dtype != torch.float64:
        raise TypeError(f"{name} must be float64 (got {x.dtype})")
    if not x.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous")
    return x


def bta_shapes(D, L, R, T):
    """Validate the BTA block arrays; return (n, b, a, device)."""
    _tensor(D, "diagonal blocks")
    if D.dim() != 3 or D.shape[1] != D.shape[2]:
        raise ValueError(f"diagonal blocks must have shape (n, b, b), got {tuple(D.shape)}")
    n, b = D.shape[0], D.shape[1]
    if n < 1 or b < 1:
        raise ValueError("need at least one diagonal block of size >= 1")
    _tensor(L, "lower diagonal blocks")
    if tuple(L.shape) != (n - 1, b, b):
        raise ValueError(f"lower diagonal blocks must have shape {(n - 1, b, b)}, got {tuple(L.shape)}")
    if R is None:
        a = 0
    else:
        _tensor(R, "lower arrow blocks")
        if R.dim() != 3 or R.shape[0] != n or R.shape[2] != b:
            raise ValueError(f"lower arrow blocks must have shape (n, a, b) = ({n}, a, {b}), got {tuple(R.shape)}")
        a = R.shape[1]
        _tensor(T, "arrow tip block")
        if tuple(T.shape) != (a, a):
            raise ValueError(f"arrow tip block must have shape {(a, a)}, got {tuple(T.shape)}")
    devs = {t.device for t in (D, L, R, T) if t is not None}
    if len(devs) != 1:
        raise ValueError(f"all blocks must be on the same device, got {devs}")
    return n, b, a, D.device


def ptr(x):
    return None if x is None or x.numel() == 0 else x.data_ptr()


def stream_of(device):
    return torch.cuda.current_stream(device).cuda_stream


def workspace(op, n, b, a, r, device):
    nbytes = _lib.load().rocbta_workspace_bytes(op, n, b, a, r)
    if nbytes == 0:
        raise ValueError(f"unsupported shape n={n} b={b} a={a}")
    return torch.empty(nbytes, dtype=torch.uint8, device=device), nbytes


def raise_on_info(info, what):
    if not config["check_info"]:
        return
    v = int(info.item())
    if v < 0:
        raise _lib.RocbtaError(f"{what}: internal scheduling time-out in the factorization kernel")
    if v != 0:
        # numpy/scipy Cholesky raise numpy.linalg.LinAlgError (the reference's
        # host path, src/serinv/__init__.py:151)
        raise np.linalg.LinAlgError(
            f"{what}: the leading minor of order {v} is not positive definite")
