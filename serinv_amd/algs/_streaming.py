"""Host <-> device streaming forms (device_streaming=True).

Reference: src/serinv/algs/pobtaf.py:284-524 (_pobtaf_streaming),
pobtasi.py:247-496 (_pobtasi_streaming), pobtf.py:188-329 and
pobtsi.py:182-330 (the arrowless forms).  As there, the block arrays are
HOST arrays (here: float64 CPU tensors, best page-locked with
``pin_memory()``), overwritten in place; the GPU holds two slots per block
kind and the recurrence streams through them (rocbta_dpobtaf_streaming /
rocbta_dpobtasi_streaming).  The calls return when the host arrays hold the
results.
"""

import numpy as np
import torch

from .. import _lib
from ._common import config, raise_on_info, stream_of

# src/serinv/algs/pobtaf.py:292-295 (the same check in the other routines)
_HOST_ONLY = "Host<->Device streaming only works when host-arrays are given."


def _host(x, name):
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor, got {type(x).__name__}")
    if x.device.type != "cpu":
        raise NotImplementedError(_HOST_ONLY)
    if x.dtype != torch.float64:
        raise TypeError(f"{name} must be float64 (got {x.dtype})")
    if not x.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous")
    return x


def host_shapes(D, L, R, T):
    """Validate host BTA block arrays; return (n, b, a)."""
    _host(D, "diagonal blocks")
    if D.dim() != 3 or D.shape[1] != D.shape[2]:
        raise ValueError(f"diagonal blocks must have shape (n, b, b), got {tuple(D.shape)}")
    n, b = D.shape[0], D.shape[1]
    if n < 1 or b < 1:
        raise ValueError("need at least one diagonal block of size >= 1")
    _host(L, "lower diagonal blocks")
    if tuple(L.shape) != (n - 1, b, b):
        raise ValueError(f"lower diagonal blocks must have shape {(n - 1, b, b)}, got {tuple(L.shape)}")
    a = 0
    if R is not None:
        _host(R, "lower arrow blocks")
        if R.dim() != 3 or R.shape[0] != n or R.shape[2] != b:
            raise ValueError(f"lower arrow blocks must have shape (n, a, {b}), got {tuple(R.shape)}")
        a = R.shape[1]
        _host(T, "arrow tip block")
        if tuple(T.shape) != (a, a):
            raise ValueError(f"arrow tip block must have shape {(a, a)}, got {tuple(T.shape)}")
    return n, b, a


def _ptr(x):
    return None if x is None or x.numel() == 0 else x.data_ptr()


def _device():
    if not torch.cuda.is_available():
        raise _lib.RocbtaError("device streaming needs a visible GPU")
    return torch.device("cuda", torch.cuda.current_device())


def factor_streaming(D, L, R, T, n, b, a, factorize_last_block, what):
    lib = _lib.load()
    dev = _device()
    nbytes = lib.rocbta_workspace_bytes(_lib.OP_POBTAF_STREAMING, n, b, a, 0)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    info = torch.empty(1, dtype=torch.int32, device=dev)
    rc = lib.rocbta_dpobtaf_streaming(n, b, a, _ptr(D), _ptr(L), _ptr(R), _ptr(T),
                                      int(factorize_last_block), info.data_ptr(), ws.data_ptr(),
                                      nbytes, stream_of(dev))
    _lib.check(rc, what)
    if config["check_info"]:
        raise_on_info(info, what)
    elif int(info.item()) > 0:  # the call is synchronous anyway: report like a checked call
        raise np.linalg.LinAlgError(
            f"{what}: the leading minor of order {int(info.item())} is not positive definite")


def selinv_streaming(D, L, R, T, n, b, a, invert_last_block, what):
    lib = _lib.load()
    dev = _device()
    nbytes = lib.rocbta_workspace_bytes(_lib.OP_POBTASI_STREAMING, n, b, a, 0)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    rc = lib.rocbta_dpobtasi_streaming(n, b, a, _ptr(D), _ptr(L), _ptr(R), _ptr(T),
                                       int(invert_last_block), ws.data_ptr(), nbytes,
                                       stream_of(dev))
    _lib.check(rc, what)
