"""ctypes binding of librocbta.so (the C ABI declared in include/rocbta.h).

The library is loaded lazily on first use.  There is no fallback: if the
library is missing the call raises, so nothing silently runs elsewhere.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "librocbta.so")

ROCBTA_OK = 0
ROCBTA_ERR_ARG = -1
ROCBTA_ERR_HIP = -2
ROCBTA_ERR_WORKSPACE = -3
OP_POBTAF, OP_POBTASI, OP_POBTAS, OP_RESIDUAL, OP_LOGDET = 0, 1, 2, 3, 4
OP_POBTAF_PERMUTED, OP_POBTASI_PERMUTED, OP_POBTAS_PERMUTED = 5, 6, 7
OP_POBTAF_STREAMING, OP_POBTASI_STREAMING = 8, 9
OP_DDBTASC, OP_DDBTASCI = 10, 11
BATCHED_FACTOR, BATCHED_SOLVE, BATCHED_SELINV, BATCHED_MIXED = 1, 2, 4, 8
BATCHED_RC = {1: 16, 4: 32, 8: 64}  # forced sweep column-chunk width
BATCHED_GEMM512 = 128

_ERRORS = {
    ROCBTA_ERR_ARG: "invalid argument",
    ROCBTA_ERR_HIP: "HIP runtime error",
    ROCBTA_ERR_WORKSPACE: "workspace too small",
}

_c_i64 = ctypes.c_int64
_c_p = ctypes.c_void_p
_c_int = ctypes.c_int
_c_size = ctypes.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "rocbta_version": (ctypes.c_char_p, []),
    "rocbta_workspace_bytes": (_c_size, [_c_int, _c_i64, _c_i64, _c_i64, _c_i64]),
    "rocbta_dpobtaf": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_int, _c_p,
                                _c_p, _c_size, _c_p]),
    "rocbta_noalloc_scratch_bytes": (_c_size, []),
    "rocbta_dpobtaf_noalloc": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_int, _c_p,
                                        _c_p]),
    "rocbta_dpobtasi": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_int, _c_p,
                                 _c_size, _c_p]),
    "rocbta_dpobtas": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i64,
                                _c_i64, ctypes.c_char, _c_int, _c_p, _c_size, _c_p]),
    "rocbta_dpobtaf_permuted": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                         _c_p, _c_size, _c_p]),
    "rocbta_dpobtasi_permuted": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                          _c_size, _c_p]),
    "rocbta_dpobtas_permuted": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                         _c_i64, _c_i64, ctypes.c_char, _c_p, _c_size, _c_p]),
    "rocbta_dpobtaf_streaming": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_int, _c_p,
                                          _c_p, _c_size, _c_p]),
    "rocbta_dpobtasi_streaming": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_int,
                                           _c_p, _c_size, _c_p]),
    "rocbta_dddbtasc": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
                                 _c_int, _c_p, _c_size, _c_p]),
    "rocbta_dddbtasci": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
                                  _c_int, _c_p, _c_size, _c_p]),
    "rocbta_dddbtasc_quadratic": (_c_int, [_c_i64, _c_i64, _c_i64] + [_c_p] * 12 +
                                  [_c_int, _c_int, _c_p, _c_size, _c_p]),
    "rocbta_dddbtasci_quadratic": (_c_int, [_c_i64, _c_i64, _c_i64] + [_c_p] * 12 +
                                   [_c_int, _c_int, _c_p, _c_size, _c_p]),
    "rocbta_dddbtasc_permuted": (_c_int, [_c_i64, _c_i64, _c_i64] + [_c_p] * 8 +
                                 [_c_p, _c_size, _c_p]),
    "rocbta_dddbtasci_permuted": (_c_int, [_c_i64, _c_i64, _c_i64] + [_c_p] * 8 +
                                  [_c_p, _c_size, _c_p]),
    "rocbta_dddbtasc_quadratic_permuted": (_c_int, [_c_i64, _c_i64, _c_i64] + [_c_p] * 16 +
                                           [_c_p, _c_size, _c_p]),
    "rocbta_dddbtasci_quadratic_permuted": (_c_int, [_c_i64, _c_i64, _c_i64] + [_c_p] * 16 +
                                            [_c_p, _c_size, _c_p]),
    "rocbta_dlogdet": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_size, _c_p]),
    "rocbta_dag_selfcheck": (_c_i64, [_c_i64, _c_i64, _c_i64]),
    "rocbta_mpobtaf": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_int, _c_p,
                                _c_p, _c_size, _c_p]),
    "rocbta_mpobtasi": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_int, _c_p,
                                 _c_size, _c_p]),
    "rocbta_dresidual": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                  _c_i64, _c_i64, _c_p, _c_size, _c_p]),
    "rocbta_dgrad": (_c_int, [_c_i64, _c_i64, _c_i64, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i64, _c_i64,
                              _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "rocbta_dquad": (_c_int, [_c_i64, _c_i64, _c_p, _c_p, _c_i64, _c_p, _c_p]),
    "rocbta_batched_create": (_c_int, [_c_int, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                       _c_int, ctypes.POINTER(_c_p)]),
    "rocbta_batched_set_grid": (_c_int, [_c_p, _c_int, _c_int]),
    "rocbta_batched_run": (_c_int, [_c_p, _c_int, _c_p]),
    "rocbta_batched_results": (_c_int, [_c_p, _c_p, _c_p, _c_p]),
    "rocbta_batched_launch_count": (_c_int, [_c_p]),
    "rocbta_batched_destroy": (_c_int, [_c_p]),
}

_lib = None
_lock = threading.Lock()


class RocbtaError(RuntimeError):
    pass


def load():
    """Load librocbta.so and declare its signatures (idempotent)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RocbtaError(
                    f"{LIB_PATH} is missing: build it first "
                    "(python -m serinv_amd.csrc.build or __graft_entry__.build())")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc, what):
    if rc != ROCBTA_OK:
        raise RocbtaError(f"{what} failed: {_ERRORS.get(rc, rc)}")
