"""Entry points called from several host threads at once, each on its own stream.

pobtasi overlaps its independent precompute on a side stream and joins it
with events; two host threads calling it concurrently must not record or wait
on each other's events.  Each thread repeats pobtaf + pobtas + pobtasi on its
own matrix and stream, and every repetition is compared with the oracle at
the FP64 parity bar.
"""

import threading

import numpy as np
import pytest

from oracle import serinv_np
from serinv_amd.utils.matrices import bta_spd
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12


def test_two_threads_two_streams(gpu_device):
    import torch

    import serinv_amd.algs as algs

    shapes = [(12, 192, 16), (9, 256, 24)]
    inputs, refs = [], []
    for k, (n, b, a) in enumerate(shapes):
        D, L, R, T, B = bta_spd(n, b, a, seed=700 + k, n_rhs=3)
        inputs.append((D, L, R, T, B))
        Do, Lo, Ro, To, Bo = (x.copy() for x in (D, L, R, T, B))
        serinv_np.pobtaf(Do, Lo, Ro, To)
        serinv_np.pobtas(Do, Lo, Ro, To, Bo, trans="N")
        serinv_np.pobtas(Do, Lo, Ro, To, Bo, trans="C")
        serinv_np.pobtasi(Do, Lo, Ro, To)
        refs.append((Do, Lo, Ro, To, Bo))
    results = [[] for _ in shapes]
    errors = []
    start = threading.Barrier(len(shapes))

    def worker(k):
        try:
            s = torch.cuda.Stream(device=gpu_device)
            with torch.cuda.stream(s):
                base = to_dev(gpu_device, *inputs[k])
                start.wait()
                outs = []
                for _ in range(6):
                    d, l, r, t, bb = (x.clone() for x in base)
                    algs.pobtaf(d, l, r, t)
                    algs.pobtas(d, l, r, t, bb, trans="N")
                    algs.pobtas(d, l, r, t, bb, trans="C")
                    algs.pobtasi(d, l, r, t)
                    outs.append((d, l, r, t, bb))
                s.synchronize()
                results[k] = [to_host(*o) for o in outs]
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(shapes))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not errors, errors
    for k, outs in enumerate(results):
        assert len(outs) == 6
        for rep, out in enumerate(outs):
            for nm, g, o in zip(("X_D", "X_L", "X_R", "X_T", "x"), out, refs[k]):
                assert relerr(g, o) < TOL, f"thread {k} repetition {rep} {nm}: {relerr(g, o):.2e}"
