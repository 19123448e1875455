"""Repeated persistent factorizations at the full workload size.

The chain pobtaf kernel hands tiles between workgroups through flags.  An
ordering fault there (a consumer reading a tile before every wave's stores
of it are visible) does not show at the small parity sizes: it shows as an
intermittent, finite error in some factor tiles at n=256, b=1024, a=32,
where about every second factorization was affected before the publish
protocol waited for the stores of all waves.  Each repetition is compared
with the launch-per-tile factor (ROCBTA_POBTAF_ALGO=loop), which has no
inter-workgroup hand-off, at the FP64 parity bar.
"""

import pytest
import torch

pytestmark = pytest.mark.gpu
TOL = 1e-12
REPEATS = 4


def test_chain_repeated_full_size(gpu_device, monkeypatch):
    import serinv_amd.algs as algs
    from serinv_amd.utils.conditioning import ShiftedBTA

    n, b, a = 256, 1024, 32
    sh = ShiftedBTA(n, b, a, gpu_device)
    A0 = sh.blocks(0.1650719305645498)
    monkeypatch.setenv("ROCBTA_POBTAF_ALGO", "loop")
    ref = [x.clone() for x in A0]
    algs.pobtaf(*ref)
    monkeypatch.delenv("ROCBTA_POBTAF_ALGO")
    F = [torch.empty_like(x) for x in A0]
    for rep in range(REPEATS):
        for f, x in zip(F, A0):
            f.copy_(x)
        algs.pobtaf(*F)
        for name, f, r in zip("DLRT", F, ref):
            # max |f - r| / max |r| over the array (the parity-test measure)
            err = float((f - r).abs().max() / r.abs().max())
            assert err < TOL, f"repetition {rep}: factor {name} differs from the loop factor ({err:.2e})"
