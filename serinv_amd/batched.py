"""Batched SPD BTA systems: one plan, replayed (no counterpart in the reference).

    plan = BatchedBTA([(D0, L0, R0, T0), (D1, L1, R1, T1), ...],
                      rhs=[B0, B1, ...], selinv=True)
    plan.run()                       # in place, like pobtaf / pobtas / pobtasi
    info, logdet, status = plan.results()

Each system uses serinv's block layout (float64 ROCm tensors; a <= 64); the
systems may differ in n, b, a and the number of right-hand sides.  A run
factors every matrix in one persistent launch, computes the
log-determinants, overwrites each B with A^{-1} B (one persistent launch per
sweep direction) and, with selinv=True, overwrites each factor with the
selected inverse (merged grouped-GEMM phases).  With graph=True the launch
sequence is captured into a hipGraph on the first run and replayed after.
With mixed=True the factorization and the selected inversion run in the
mixed FP32/FP64 form of serinv_amd.mixed (FP32 tile products; FP64 tile
potrf/trtri, tip, log-determinant and solves).

Launch shapes (picked per batch class by serinv_amd.autotune): sweep_width
forces the solve sweeps' column-chunk width (1, 4 or 8 right-hand-side
columns per chain workgroup; None: 1 for one column, 4 up to four, else 8),
gemm_threads (256 or 512) the workgroup size of the selected-inversion GEMM
phases (512: two halves split each 64 x 64 tile's K range).
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .algs._common import _tensor, bta_shapes


class BatchedBTA:
    def __init__(self, systems, rhs=None, selinv=True, graph=True, mixed=False,
                 sweep_width=None, gemm_threads=None):
        if not systems:
            raise ValueError("need at least one system")
        if sweep_width not in (None, 1, 4, 8):
            raise ValueError(f"sweep_width must be 1, 4 or 8 (got {sweep_width})")
        if gemm_threads not in (None, 256, 512):
            raise ValueError(f"gemm_threads must be 256 or 512 (got {gemm_threads})")
        self._keep = []
        ns, bs, as_, rs = [], [], [], []
        ptrs = {k: [] for k in "DLRTB"}
        dev = None
        for k, (D, L, R, T) in enumerate(systems):
            n, b, a, d = bta_shapes(D, L, R, T)
            if a > 64:
                raise ValueError(f"system {k}: the batched plan supports a <= 64 (got {a})")
            if dev is None:
                dev = d
            elif d != dev:
                raise ValueError("all systems must be on the same device")
            ns.append(n)
            bs.append(b)
            as_.append(a)
            for key, t in zip("DLRT", (D, L, R, T)):
                ptrs[key].append(t.data_ptr() if t is not None and t.numel() else 0)
            self._keep += [D, L, R, T]
            if rhs is not None:
                B = _tensor(rhs[k], f"rhs[{k}]")
                if B.dim() == 1:
                    r = 1
                elif B.dim() == 2:
                    r = B.shape[1]
                else:
                    raise ValueError(f"rhs[{k}] must be 1-D or 2-D")
                if B.shape[0] != n * b + a or B.device != dev:
                    raise ValueError(f"rhs[{k}] must have {n * b + a} rows on the systems' device")
                if r > 512:
                    raise ValueError(f"rhs[{k}]: the batched plan solves at most 512 columns per system "
                                     f"(got {r}); solve wider right-hand sides with pobtas")
                rs.append(r)
                ptrs["B"].append(B.data_ptr())
                self._keep.append(B)
        nmat = len(systems)
        self.nmat = nmat
        self.device = dev
        self.graph = bool(graph)
        ops = _lib.BATCHED_FACTOR
        if rhs is not None:
            if len(rhs) != nmat:
                raise ValueError("rhs must have one entry per system")
            ops |= _lib.BATCHED_SOLVE
        if selinv:
            ops |= _lib.BATCHED_SELINV
        if mixed:
            ops |= _lib.BATCHED_MIXED
        if sweep_width is not None:
            ops |= _lib.BATCHED_RC[sweep_width]
        if gemm_threads == 512:
            ops |= _lib.BATCHED_GEMM512
        I64 = ctypes.c_int64 * nmat
        P = ctypes.c_void_p * nmat
        self._arrays = [I64(*ns), I64(*bs), I64(*as_), I64(*(rs or [0] * nmat))]
        self._ptrs = {k: P(*(v or [0] * nmat)) for k, v in ptrs.items()}
        lib = _lib.load()
        handle = ctypes.c_void_p()
        with torch.cuda.device(dev):
            vp = lambda x: ctypes.cast(x, ctypes.c_void_p)  # noqa: E731
            rc = lib.rocbta_batched_create(
                nmat, vp(self._arrays[0]), vp(self._arrays[1]), vp(self._arrays[2]),
                vp(self._ptrs["D"]), vp(self._ptrs["L"]), vp(self._ptrs["R"]), vp(self._ptrs["T"]),
                vp(self._ptrs["B"]) if rhs is not None else None,
                vp(self._arrays[3]) if rhs is not None else None, ops, ctypes.byref(handle))
        _lib.check(rc, "BatchedBTA")
        self._plan = handle
        # a capturable stream of its own (graph capture cannot use the null stream)
        self._stream = torch.cuda.Stream(device=dev)

    def set_grid(self, factor_grid=0, solve_grid=0):
        """Workgroup counts of the persistent launches (0 keeps the current)."""
        _lib.check(_lib.load().rocbta_batched_set_grid(self._plan, int(factor_grid), int(solve_grid)),
                   "BatchedBTA.set_grid")

    def run(self):
        cur = torch.cuda.current_stream(self.device)
        self._stream.wait_stream(cur)
        with torch.cuda.device(self.device):
            rc = _lib.load().rocbta_batched_run(self._plan, int(self.graph), self._stream.cuda_stream)
        cur.wait_stream(self._stream)
        _lib.check(rc, "BatchedBTA.run")

    def results(self):
        """Synchronise; return (info int array, logdet float64 array, solve status)."""
        info = np.zeros(self.nmat, dtype=np.int32)
        logdet = np.zeros(self.nmat, dtype=np.float64)
        status = ctypes.c_int(0)
        with torch.cuda.device(self.device):
            rc = _lib.load().rocbta_batched_results(
                self._plan, info.ctypes.data_as(ctypes.c_void_p),
                logdet.ctypes.data_as(ctypes.c_void_p), ctypes.byref(status))
        _lib.check(rc, "BatchedBTA.results")
        return info, logdet, status.value

    def launch_count(self):
        return _lib.load().rocbta_batched_launch_count(self._plan)

    def close(self):
        if getattr(self, "_plan", None):
            torch.cuda.synchronize(self.device)
            _lib.load().rocbta_batched_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
