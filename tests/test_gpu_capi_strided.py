"""rocbta_dpobtas called through the C-ABI on a column range of a wider
right-hand-side array (row stride ldb > r, pointer at a column offset), for
each solve algorithm: the block sweep, its wide form, the tile sweep and
the launch-per-tile form (ROCBTA_POBTAS_ALGO), with the workspace filled
with NaN beforehand (the solves read only workspace they wrote).  Checked
against the oracle's pobtas (src/serinv/algs/pobtas.py:65-161) on those
columns at 1e-12; the columns outside the range must come back untouched.
"""

import ctypes

import numpy as np
import pytest

from oracle import serinv_np
from serinv_amd.utils.matrices import bta_spd
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("algo,n,b,a,r", [("block", 3, 256, 8, 3), ("wide", 4, 64, 5, 70),
                                          ("wide", 3, 256, 9, 65), ("block", 2, 128, 0, 1),
                                          ("sweep", 3, 70, 4, 9), ("loop", 2, 40, 3, 5)])
def test_pobtas_strided_columns(gpu_device, monkeypatch, algo, n, b, a, r):
    import torch

    from serinv_amd import _lib
    from serinv_amd.algs._common import ptr, stream_of

    monkeypatch.setenv("ROCBTA_POBTAS_ALGO", algo)
    print(f"strided pobtas {algo} n={n} b={b} a={a} r={r}", flush=True)
    lib = _lib.load()
    D, L, R, T, _ = bta_spd(n, b, a, seed=1400 + b + r, n_rhs=1)
    serinv_np.pobtaf(D, L, R, T)
    ldb, c0 = r + 23, 11
    Bfull = np.random.default_rng(r).uniform(-1, 1, (n * b + a, ldb))
    d, l, rr, t, bf = to_dev(gpu_device, D, L, R, T, Bfull)
    nbytes = lib.rocbta_workspace_bytes(_lib.OP_POBTAS, n, b, a, r)
    # workspace poisoned with NaN: a read of anything the solve did not write
    # first turns the result into NaN
    ws = torch.empty(max(nbytes, 8) // 8 * 8 + 8, dtype=torch.uint8, device=gpu_device)
    ws.view(torch.float64).fill_(float("nan"))
    ref = Bfull.copy()
    sub = ref[:, c0:c0 + r].copy()
    for trans in (b"N", b"C"):
        serinv_np.pobtas(D, L, R, T, sub, trans=trans.decode())
        rc = lib.rocbta_dpobtas(n, b, a, ptr(d), ptr(l), ptr(rr), ptr(t),
                                bf.data_ptr() + c0 * 8, r, ldb, trans, 0, ws.data_ptr(), nbytes,
                                stream_of(gpu_device))
        assert rc == 0
    torch.cuda.synchronize()
    assert int(ws[:4].view(torch.int32).item()) == 0  # no scheduling time-out
    (got,) = to_host(bf)
    assert relerr(got[:, c0:c0 + r], sub) < TOL
    assert np.array_equal(got[:, :c0], Bfull[:, :c0])
    assert np.array_equal(got[:, c0 + r:], Bfull[:, c0 + r:])


@pytest.mark.parametrize("n,b,a", [(4, 256, 8), (3, 128, 0), (3, 70, 5)])
def test_pobtasi_poisoned_workspace(gpu_device, n, b, a):
    """rocbta_dpobtasi through the C-ABI with its workspace filled with NaN
    beforehand, against the oracle's pobtasi (src/serinv/algs/pobtasi.py:
    92-172) at 1e-12: the selected inversion reads only workspace it wrote."""
    import torch

    from serinv_amd import _lib
    from serinv_amd.algs._common import ptr, stream_of

    print(f"poisoned pobtasi n={n} b={b} a={a}", flush=True)
    lib = _lib.load()
    D, L, R, T, _ = bta_spd(n, b, a, seed=1500 + b, n_rhs=1)
    serinv_np.pobtaf(D, L, R, T)
    d, l, rr, t = to_dev(gpu_device, D, L, R, T)
    nbytes = lib.rocbta_workspace_bytes(_lib.OP_POBTASI, n, b, a, 0)
    ws = torch.empty(nbytes // 8 * 8 + 8, dtype=torch.uint8, device=gpu_device)
    ws.view(torch.float64).fill_(float("nan"))
    assert lib.rocbta_dpobtasi(n, b, a, ptr(d), ptr(l), ptr(rr), ptr(t), 1, ws.data_ptr(), nbytes,
                               stream_of(gpu_device)) == 0
    torch.cuda.synchronize()
    serinv_np.pobtasi(D, L, R, T)
    for g, o in zip(to_host(d, l, rr, t), (D, L, R, T)):
        assert relerr(g, o) < TOL
