"""GPU checks of the batched plan (serinv_amd.batched.BatchedBTA).

Every system of an uneven batch is checked against the oracle run on that
system alone (factor via logdet, solve, selected inverse) to the FP64 bar
1e-12, and a captured hipGraph is replayed several times, with the inputs
restored by eager copies in between, each replay matching the eager run
bit for bit.
"""

import numpy as np
import pytest

from oracle import serinv_np
from serinv_amd.utils.matrices import bta_spd
from tests._helpers import relerr, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12

SHAPES = [(3, 64, 16, 4), (1, 8, 5, 1), (5, 70, 0, 3), (2, 128, 64, 9), (4, 1, 1, 2), (6, 100, 33, 4),
          (2, 192, 7, 1)]


def _oracle(D, L, R, T, B):
    D, L, R, T, B = (x.copy() for x in (D, L, R, T, B))
    serinv_np.pobtaf(D, L, R, T)
    ld = serinv_np.logdet_from_factor(D, T)
    serinv_np.pobtas(D, L, R, T, B, trans="N")
    serinv_np.pobtas(D, L, R, T, B, trans="C")
    serinv_np.pobtasi(D, L, R, T)
    return ld, B, (D, L, R, T)


def _make(gpu_device, shapes, seed0=300):
    host, dev = [], []
    for k, (n, b, a, r) in enumerate(shapes):
        D, L, R, T, B = bta_spd(n, b, a, seed=seed0 + k, n_rhs=r)
        host.append((D, L, R, T, B))
        dev.append(to_dev(gpu_device, D, L, R, T, B))
    return host, dev


@pytest.mark.parametrize("graph", [False, True])
def test_batched_matches_oracle(gpu_device, graph):
    from serinv_amd.batched import BatchedBTA

    host, dev = _make(gpu_device, SHAPES)
    plan = BatchedBTA([d[:4] for d in dev], rhs=[d[4] for d in dev], selinv=True, graph=graph)
    plan.run()
    info, logdet, status = plan.results()
    assert status == 0 and np.all(info == 0), (info, status)
    for k, (h, d) in enumerate(zip(host, dev)):
        ld, x, X = _oracle(*h)
        assert abs(logdet[k] - ld) <= TOL * abs(ld), f"system {k} logdet"
        assert relerr(to_host(d[4])[0], x) < TOL, f"system {k} solve {relerr(to_host(d[4])[0], x):.2e}"
        for nm, g, o in zip("DLRT", to_host(*d[:4]), X):
            assert relerr(g, o) < TOL, f"system {k} X_{nm}: {relerr(g, o):.2e}"
    plan.close()


def test_batched_graph_replays(gpu_device):
    import torch

    from serinv_amd.batched import BatchedBTA

    host, dev = _make(gpu_device, SHAPES[:4], seed0=400)
    orig = [[t.clone() for t in d] for d in dev]
    eager = BatchedBTA([d[:4] for d in dev], rhs=[d[4] for d in dev], graph=False)
    eager.run()
    ref = [[t.clone() for t in d] for d in dev]
    ref_ld = eager.results()[1]
    eager.close()
    plan = BatchedBTA([d[:4] for d in dev], rhs=[d[4] for d in dev], graph=True)
    for rep in range(3):
        for d, o in zip(dev, orig):
            for t, s in zip(d, o):
                t.copy_(s)
        torch.cuda.synchronize()
        plan.run()
        info, ld, status = plan.results()
        assert status == 0 and np.all(info == 0)
        assert np.array_equal(ld, ref_ld), f"replay {rep} logdet"
        for d, rr in zip(dev, ref):
            for t, s in zip(d, rr):
                assert torch.equal(t, s), f"replay {rep} differs from the eager run"
    plan.close()


def test_batched_reports_non_pd(gpu_device):
    from serinv_amd.batched import BatchedBTA

    host, dev = _make(gpu_device, [(3, 16, 4, 1), (2, 8, 2, 1)], seed0=500)
    dev[1][0][1].copy_(-dev[1][0][1])  # second system: D_1 negative definite
    plan = BatchedBTA([d[:4] for d in dev], rhs=[d[4] for d in dev], selinv=False, graph=False)
    plan.run()
    info, _, _ = plan.results()
    assert info[0] == 0 and info[1] == 9, info  # first pivot of block 1 = global row 9
    plan.close()


def test_batched_more_systems_than_cus(gpu_device):
    """More systems than CUs: the factorization chains and the sweep chains
    run in waves, so helpers always find CUs; every system is checked."""
    import torch

    from serinv_amd.batched import BatchedBTA

    nmat = torch.cuda.get_device_properties(gpu_device).multi_processor_count + 44
    shapes = [(2 + k % 3, 16 + 8 * (k % 4), 1 + k % 5, 1 + k % 3) for k in range(nmat)]
    host, dev = _make(gpu_device, shapes, seed0=1000)
    plan = BatchedBTA([d[:4] for d in dev], rhs=[d[4] for d in dev], selinv=True, graph=True)
    plan.run()
    info, logdet, status = plan.results()
    plan.close()
    assert status == 0 and np.all(info == 0), (info, status)
    for k, (h, d) in enumerate(zip(host, dev)):
        ld, x, X = _oracle(*h)
        assert abs(logdet[k] - ld) <= TOL * abs(ld), f"system {k} logdet"
        assert relerr(to_host(d[4])[0], x) < TOL, f"system {k} solve"
        for nm, g, o in zip("DLRT", to_host(*d[:4]), X):
            assert relerr(g, o) < TOL, f"system {k} X_{nm}: {relerr(g, o):.2e}"


def test_batched_mixed(gpu_device, monkeypatch):
    """Mixed FP32/FP64 plan: every system against the FP64 oracle at the bar
    of the mixed form (FP32 products, 1e-5), and its factor equal to the
    per-matrix mixed factorization with the step-after-step chain schedule
    (ROCBTA_CHAIN=classic), which runs the same tile kernels as the plan."""
    from serinv_amd.batched import BatchedBTA
    from serinv_amd.logdet import pobta_logdet
    from serinv_amd.mixed import pobtaf_mixed

    host, dev = _make(gpu_device, SHAPES, seed0=1500)
    single = []
    monkeypatch.setenv("ROCBTA_CHAIN", "classic")
    for d in dev:
        f = [t.clone() for t in d[:4]]
        pobtaf_mixed(*f)
        single.append((f, float(pobta_logdet(f[0], f[3]))))
    monkeypatch.delenv("ROCBTA_CHAIN")
    plan = BatchedBTA([d[:4] for d in dev], rhs=[d[4] for d in dev], selinv=False, graph=True,
                      mixed=True)
    plan.run()
    info, logdet, status = plan.results()
    plan.close()
    assert status == 0 and np.all(info == 0), (info, status)
    for k, (h, d) in enumerate(zip(host, dev)):
        D, L, R, T, B = (x.copy() for x in h)
        serinv_np.pobtaf(D, L, R, T)
        ld = serinv_np.logdet_from_factor(D, T)
        for nm, g, s, o in zip("DLRT", to_host(*d[:4]), to_host(*single[k][0]), (D, L, R, T)):
            assert np.array_equal(g, s), f"system {k} factor {nm} differs from pobtaf_mixed"
            assert relerr(g, o) < 1e-5, f"system {k} factor {nm}: {relerr(g, o):.2e}"
        assert logdet[k] == single[k][1] or abs(logdet[k] - single[k][1]) <= 1e-14 * abs(ld)
        assert abs(logdet[k] - ld) <= 1e-7 * abs(ld)
        serinv_np.pobtas(D, L, R, T, B, trans="N")
        serinv_np.pobtas(D, L, R, T, B, trans="C")
        assert relerr(to_host(d[4])[0], B) < 1e-5, f"system {k} solve"
    # selected inversion in the mixed plan
    host, dev = _make(gpu_device, SHAPES[:4], seed0=1600)
    plan = BatchedBTA([d[:4] for d in dev], selinv=True, graph=False, mixed=True)
    plan.run()
    info, _, _ = plan.results()
    plan.close()
    assert np.all(info == 0)
    for k, (h, d) in enumerate(zip(host, dev)):
        _, _, X = _oracle(*h)
        for nm, g, o in zip("DLRT", to_host(*d[:4]), X):
            assert relerr(g, o) < 1e-5, f"system {k} X_{nm}: {relerr(g, o):.2e}"


@pytest.mark.parametrize("sweep_width,gemm_threads", [(1, 512), (8, 256), (4, 512)])
def test_batched_launch_shapes(gpu_device, sweep_width, gemm_threads):
    """Forced tile shapes (sweep column-chunk width, GEMM workgroup size)
    against the oracle at the FP64 bar."""
    from serinv_amd.batched import BatchedBTA

    host, dev = _make(gpu_device, SHAPES, seed0=1700)
    plan = BatchedBTA([d[:4] for d in dev], rhs=[d[4] for d in dev], selinv=True, graph=True,
                      sweep_width=sweep_width, gemm_threads=gemm_threads)
    plan.run()
    info, logdet, status = plan.results()
    plan.close()
    assert status == 0 and np.all(info == 0), (info, status)
    for k, (h, d) in enumerate(zip(host, dev)):
        ld, x, X = _oracle(*h)
        assert abs(logdet[k] - ld) <= TOL * abs(ld), f"system {k} logdet"
        assert relerr(to_host(d[4])[0], x) < TOL, f"system {k} solve"
        for nm, g, o in zip("DLRT", to_host(*d[:4]), X):
            assert relerr(g, o) < TOL, f"system {k} X_{nm}: {relerr(g, o):.2e}"


def test_autotuner_round_trip(gpu_device):
    """tune_batched times grids then tile shapes; the tuned plan matches the
    oracle."""
    import torch

    from serinv_amd.autotune import apply_tuning, tune_batched

    host, dev = _make(gpu_device, SHAPES[:4], seed0=1800)
    cus = torch.cuda.get_device_properties(gpu_device).multi_processor_count
    tuned = tune_batched([d[:4] for d in dev], rhs=[d[4] for d in dev], selinv=True, reps=1,
                         use_cache=False, candidates=[(cus, cus), (2 * cus, cus)])
    assert tuned["sweep_width"] in (None, 1, 4, 8) and tuned["gemm_threads"] in (None, 256, 512)
    shapes = {(t["sweep_width"], t["gemm_threads"]) for t in tuned["trials"]}
    assert (4, 512) in shapes and (1, 256) in shapes  # r = 9 max: widths 1, 4, 8
    plan = apply_tuning([d[:4] for d in dev], tuned, rhs=[d[4] for d in dev], selinv=True)
    plan.run()
    info, logdet, status = plan.results()
    plan.close()
    assert status == 0 and np.all(info == 0)
    for k, (h, d) in enumerate(zip(host, dev)):
        ld, x, X = _oracle(*h)
        assert abs(logdet[k] - ld) <= TOL * abs(ld)
        assert relerr(to_host(d[4])[0], x) < TOL
        for g, o in zip(to_host(*d[:4]), X):
            assert relerr(g, o) < TOL
