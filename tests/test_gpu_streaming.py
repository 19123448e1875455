"""GPU parity of the host-streaming forms (device_streaming=True).

rocbta_dpobtaf_streaming / rocbta_dpobtasi_streaming, through
serinv_amd.algs with HOST tensors, against the oracle's restatement of the
natural routines (src/serinv/algs/pobtaf.py:97-183, pobtasi.py:92-172: the
streaming forms pobtaf.py:284-524 and pobtasi.py:247-496 compute the same
recurrences with the blocks streamed through the device).  Tolerance 1e-12
relative (max-norm per output array).  Shapes: n = 1 and 2 (no loop / one
step), tile boundaries (b 63, 65, 100, 130; a 1, 64, 65), the arrowless
forms, n = 16 with b = 256, a = 64; page-locked and pageable host memory;
the partial flags; a non-positive pivot.
"""

import numpy as np
import pytest

from oracle import dense, generators, serinv_np
from tests._helpers import relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def algs(gpu_device):
    import serinv_amd.algs as A

    return A


def case(b, a, n, seed):
    rng = np.random.default_rng(seed)
    A = generators.dd_bta(b, a, n, rng)
    generators.symmetrize(A)
    return A


def host(pinned, *arrays):
    import torch

    out = []
    for x in arrays:
        t = torch.from_numpy(np.ascontiguousarray(x)).clone()
        out.append(t.pin_memory() if pinned else t)
    return out


SHAPES = [(3, 2, 1), (3, 2, 2), (64, 16, 4), (100, 3, 5), (130, 65, 3), (63, 1, 4), (65, 64, 3),
          (256, 64, 16)]


@pytest.mark.parametrize("b,a,n", SHAPES)
def test_streaming_factor_and_selinv_match_oracle(algs, gpu_device, b, a, n):
    print(f"streaming b={b} a={a} n={n}", flush=True)
    A = case(b, a, n, seed=21000 + b + 7 * a + 13 * n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    hD, hL, hR, hT = host(True, D, L, R, T)
    algs.pobtaf(hD, hL, hR, hT, device_streaming=True)
    serinv_np.pobtaf(D, L, R, T)
    for x, y in ((hD, D), (hL, L), (hR, R), (hT, T)):
        assert relerr(x.numpy(), y) < TOL
    assert np.all(np.triu(hD.numpy(), 1) == 0)
    algs.pobtasi(hD, hL, hR, hT, device_streaming=True)
    serinv_np.pobtasi(D, L, R, T)
    for x, y in ((hD, D), (hL, L), (hR, R), (hT, T)):
        assert relerr(x.numpy(), y) < TOL


@pytest.mark.parametrize("b,a,n", [(64, 8, 4), (100, 3, 3)])
def test_streaming_partial_flags(algs, gpu_device, b, a, n):
    """factorize_last_block=False leaves the Schur complements (full blocks);
    invert_last_block=False takes the last blocks as already inverted."""
    print(f"streaming partial b={b} a={a} n={n}", flush=True)
    A = case(b, a, n, seed=22000 + b + a + n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    hD, hL, hR, hT = host(True, D, L, R, T)
    algs.pobtaf(hD, hL, hR, hT, device_streaming=True, factorize_last_block=False)
    serinv_np.pobtaf(D, L, R, T, factorize_last_block=False)
    for x, y in ((hD, D), (hL, L), (hR, R), (hT, T)):
        assert relerr(x.numpy(), y) < TOL
    # finish the factor, then feed the true inverse blocks of the last block row
    X = np.linalg.inv(A)
    Xd, _, Xr, Xt = dense.bta_dense_to_arrays(X, b, a, n)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    serinv_np.pobtaf(D, L, R, T)
    D[-1], R[-1], T[:] = Xd[-1], Xr[-1], Xt
    hD, hL, hR, hT = host(True, D, L, R, T)
    algs.pobtasi(hD, hL, hR, hT, device_streaming=True, invert_last_block=False)
    serinv_np.pobtasi(D, L, R, T, invert_last_block=False)
    for x, y in ((hD, D), (hL, L), (hR, R), (hT, T)):
        assert relerr(x.numpy(), y) < TOL


@pytest.mark.parametrize("b,n", [(3, 4), (64, 5), (130, 2)])
def test_streaming_bt_forms(algs, gpu_device, b, n):
    """pobtf / pobtsi with device_streaming=True (pobtf.py:188-329,
    pobtsi.py:182-330), from pageable (not page-locked) host memory."""
    print(f"streaming bt b={b} n={n}", flush=True)
    rng = np.random.default_rng(23000 + b + n)
    A = generators.dd_bt(b, n, rng)
    generators.symmetrize(A)
    D, L = dense.bt_dense_to_arrays(A, b, n)
    hD, hL = host(False, D, L)
    algs.pobtf(hD, hL, device_streaming=True)
    serinv_np.pobtf(D, L)
    assert relerr(hD.numpy(), D) < TOL and relerr(hL.numpy(), L) < TOL
    algs.pobtsi(hD, hL, device_streaming=True)
    serinv_np.pobtsi(D, L)
    assert relerr(hD.numpy(), D) < TOL and relerr(hL.numpy(), L) < TOL


def test_streaming_errors(algs, gpu_device):
    import torch

    b, a, n = 4, 2, 3
    A = case(b, a, n, 5)
    D, L, R, T = dense.bta_dense_to_arrays(A, b, a, n)
    gD, gL, gR, gT = (torch.from_numpy(x).to(gpu_device) for x in (D, L, R, T))
    with pytest.raises(NotImplementedError, match="host-arrays"):
        algs.pobtaf(gD, gL, gR, gT, device_streaming=True)
    hD, hL, hR, hT = host(True, D, L, R, T)
    with pytest.raises(NotImplementedError, match="permuted"):
        algs.pobtaf(hD, hL, hR, hT, device_streaming=True, buffer=torch.empty_like(hD))
    # a non-positive pivot in block 1 (global row b + 1)
    D2 = D.copy()
    D2[1] = -np.eye(b)
    hD, hL, hR, hT = host(True, D2, L, R, T)
    with pytest.raises(np.linalg.LinAlgError, match=f"order {b + 1}"):
        algs.pobtaf(hD, hL, hR, hT, device_streaming=True)
