"""rocbta_dpobtaf_noalloc: pobtaf with its workspace in a static device array.

Called through the C ABI (ctypes) twice in a row, the second time with a
smaller matrix, on the caller's stream; each factor and info word must equal
those of rocbta_dpobtaf with a caller-allocated workspace (same kernels: the
factor is identical, which is within the FP64 parity bar 1e-12 by far), and
the factor is checked against the oracle.  A shape whose workspace exceeds
the static array is refused with ROCBTA_ERR_WORKSPACE.
"""

import ctypes

import numpy as np
import pytest

from oracle import serinv_np
from serinv_amd.utils.matrices import bta_spd
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu


def _call(lib, fn, arrays, dev, ws=None):
    import torch

    D, L, R, T = arrays
    n, b = D.shape[0], D.shape[1]
    a = R.shape[1]
    info = torch.full((1,), 7, dtype=torch.int32, device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    p = lambda x: x.data_ptr() if x.numel() else None  # noqa: E731
    if ws is None:
        rc = fn(n, b, a, p(D), p(L), p(R), p(T), 1, info.data_ptr(), s)
    else:
        rc = fn(n, b, a, p(D), p(L), p(R), p(T), 1, info.data_ptr(), ws.data_ptr(), ws.numel(), s)
    torch.cuda.synchronize(dev)
    return rc, int(info.item())


def test_noalloc_matches_allocating_entry_point(gpu_device):
    import torch

    from serinv_amd import _lib

    lib = _lib.load()
    cap = lib.rocbta_noalloc_scratch_bytes()
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTAF, 256, 256, 32, 0) <= cap
    for k, (n, b, a) in enumerate([(24, 256, 32), (5, 100, 7)]):  # second call smaller
        D, L, R, T, _ = bta_spd(n, b, a, seed=1300 + k)
        ref = to_dev(gpu_device, D, L, R, T)
        got = [x.clone() for x in ref]
        nbytes = lib.rocbta_workspace_bytes(_lib.OP_POBTAF, n, b, a, 0)
        assert nbytes <= cap
        ws = torch.empty(nbytes, dtype=torch.uint8, device=gpu_device)
        rc, info = _call(lib, lib.rocbta_dpobtaf, ref, gpu_device, ws)
        assert rc == 0 and info == 0
        rc, info = _call(lib, lib.rocbta_dpobtaf_noalloc, got, gpu_device)
        assert rc == 0 and info == 0, (rc, info)
        for nm, g, r in zip("DLRT", got, ref):
            assert torch.equal(g, r), f"call {k}: factor {nm} differs from rocbta_dpobtaf"
        Do, Lo, Ro, To = D.copy(), L.copy(), R.copy(), T.copy()
        serinv_np.pobtaf(Do, Lo, Ro, To)
        for nm, g, o in zip("DLRT", to_host(*got), (Do, Lo, Ro, To)):
            assert relerr(g, o) < 1e-12, f"call {k}: factor {nm} vs oracle {relerr(g, o):.2e}"


def test_noalloc_reports_pivot_and_refuses_oversize(gpu_device):
    from serinv_amd import _lib

    lib = _lib.load()
    D, L, R, T, _ = bta_spd(3, 16, 4, seed=5)
    D[1] = -np.eye(16)
    arrs = to_dev(gpu_device, D, L, R, T)
    rc, info = _call(lib, lib.rocbta_dpobtaf_noalloc, arrs, gpu_device)
    assert rc == 0 and info == 17  # first pivot of block 1 = global row 17
    # n=1024, b=1024: 16k diagonal tiles of inverse storage, far above the array
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTAF, 1024, 1024, 32, 0) > lib.rocbta_noalloc_scratch_bytes()
    rc = lib.rocbta_dpobtaf_noalloc(1024, 1024, 32, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                    ctypes.c_void_p(16), ctypes.c_void_p(16), 1, ctypes.c_void_p(16),
                                    None)
    assert rc == _lib.ROCBTA_ERR_WORKSPACE
