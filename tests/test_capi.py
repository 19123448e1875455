"""The C-ABI library loads and exports every symbol the header declares (CPU).

No compute entry point is called here; only host-side argument checks,
which return before any device work.
"""

import ctypes
import os
import re

import pytest

from serinv_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "rocbta.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rocbta_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built (run __graft_entry__.build())")
    return _lib.load()


def test_header_symbols_exported(lib):
    syms = declared_symbols()
    assert "rocbta_dpobtaf" in syms and "rocbta_dpobtasi" in syms and "rocbta_dpobtas" in syms
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in rocbta.h but not exported"
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"


def test_version(lib):
    assert b"gfx950" in lib.rocbta_version()


def test_workspace_sizes(lib):
    n, b, a = 256, 1024, 32
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTAF, n, b, a, 0) >= 64 * 64 * 8
    si = lib.rocbta_workspace_bytes(_lib.OP_POBTASI, n, b, a, 0)
    assert si >= (n - 1) * b * b * 8  # the G_i = L_{i+1,i} L_ii^{-1} blocks
    assert si < (n + 64) * b * b * 8  # L_ii^{-1} is kept for one chunk of blocks only
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTAF, 0, b, a, 0) == 0  # invalid shape
    assert lib.rocbta_workspace_bytes(99, n, b, a, 0) == 0


def test_argument_errors_return_before_launch(lib):
    # null diagonal blocks / negative sizes are rejected on the host
    assert lib.rocbta_dpobtaf(4, 8, 2, None, None, None, None, 1, None, None, 0, None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtasi(0, 8, 2, None, None, None, None, 1, None, 0, None) == _lib.ROCBTA_ERR_ARG
    fake = ctypes.c_void_p(16)
    assert lib.rocbta_dpobtas(2, 4, 1, fake, fake, fake, fake, fake, 1, 1, b"X", 0, None, 0, None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtaf(2, 4, 1, fake, fake, fake, fake, 1, None, None, 0, None) == _lib.ROCBTA_ERR_WORKSPACE


@pytest.mark.parametrize("a", [0, 1, 2, 16, 31, 32, 63, 64])
def test_pobtaf_dag_schedule_selfcheck(lib, a):
    """The persistent pobtaf kernel's task decoding and dependency lists,
    executed on the host: every wait targets an earlier task, every index
    decodes back to itself, every tile lies inside its array."""
    for n in (1, 2, 3, 7):
        for b in (1, 2, 3, 17, 63, 64, 65, 100, 128, 130, 192):
            assert lib.rocbta_dag_selfcheck(n, b, a) == 0, (n, b, a)
    assert lib.rocbta_dag_selfcheck(256, 1024, 32) == 0
    assert lib.rocbta_dag_selfcheck(4, 8, 65) == -1  # arrow wider than one tile: not this kernel


def test_batched_launch_shape_bits_checked(lib):
    """Two forced sweep widths at once are rejected before any device work."""
    n = (ctypes.c_int64 * 1)(3)
    b = (ctypes.c_int64 * 1)(64)
    a = (ctypes.c_int64 * 1)(4)
    r = (ctypes.c_int64 * 1)(2)
    fake = (ctypes.c_void_p * 1)(0x1000)  # never dereferenced: the call returns first
    plan = ctypes.c_void_p()
    ops = _lib.BATCHED_FACTOR | _lib.BATCHED_SOLVE | _lib.BATCHED_RC[1] | _lib.BATCHED_RC[4]
    rc = lib.rocbta_batched_create(1, n, b, a, fake, fake, fake, fake, fake, r, ops, ctypes.byref(plan))
    assert rc == _lib.ROCBTA_ERR_ARG and not plan.value


def test_permuted_workspace_and_argument_errors(lib):
    n, b, a = 16, 256, 64
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTAF_PERMUTED, n, b, a, 0) >= 64 * 64 * 8
    # the diagonal blocks' inverses plus the per-step copies and sums
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTASI_PERMUTED, n, b, a, 0) >= (n + 6) * b * b * 8
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTAS_PERMUTED, n, b, a, 4) >= 4
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTASI_PERMUTED, 0, b, a, 0) == 0
    fake = ctypes.c_void_p(16)
    # missing buffer / info / workspace, bad trans: rejected on the host
    assert lib.rocbta_dpobtaf_permuted(3, 4, 1, fake, fake, fake, fake, None, fake, fake, 1 << 20,
                                       None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtaf_permuted(3, 4, 1, fake, fake, fake, fake, fake, None, fake, 1 << 20,
                                       None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtaf_permuted(3, 4, 1, fake, fake, fake, fake, fake, fake, None, 0,
                                       None) == _lib.ROCBTA_ERR_WORKSPACE
    assert lib.rocbta_dpobtasi_permuted(3, 4, 1, fake, fake, fake, fake, None, fake, 1 << 20,
                                        None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtasi_permuted(3, 4, 1, fake, fake, fake, fake, fake, fake, 16,
                                        None) == _lib.ROCBTA_ERR_WORKSPACE
    assert lib.rocbta_dpobtas_permuted(3, 4, 1, fake, fake, fake, fake, fake, fake, 1, 1, b"X",
                                       None, 0, None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtas_permuted(3, 4, 1, fake, fake, fake, fake, None, fake, 1, 1, b"N",
                                       None, 0, None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtas_permuted(3, 4, 1, fake, fake, fake, fake, fake, fake, 2, 1, b"N",
                                       None, 0, None) == _lib.ROCBTA_ERR_ARG  # ldb < r


def test_streaming_workspace_and_argument_errors(lib):
    b, a = 1024, 32
    # two slots per block kind, independent of n
    f = lib.rocbta_workspace_bytes(_lib.OP_POBTAF_STREAMING, 256, b, a, 0)
    assert f >= 4 * b * b * 8 and f == lib.rocbta_workspace_bytes(_lib.OP_POBTAF_STREAMING, 4, b, a, 0)
    assert lib.rocbta_workspace_bytes(_lib.OP_POBTASI_STREAMING, 256, b, a, 0) >= 10 * b * b * 8
    fake = ctypes.c_void_p(16)
    assert lib.rocbta_dpobtaf_streaming(3, 4, 1, None, fake, fake, fake, 1, fake, fake, 1 << 20,
                                        None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtaf_streaming(3, 4, 1, fake, fake, fake, fake, 1, None, fake, 1 << 20,
                                        None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtaf_streaming(3, 4, 1, fake, fake, fake, fake, 1, fake, fake, 8,
                                        None) == _lib.ROCBTA_ERR_WORKSPACE
    assert lib.rocbta_dpobtasi_streaming(3, 4, 1, fake, fake, None, fake, 1, fake, 1 << 20,
                                         None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dpobtasi_streaming(3, 4, 1, fake, fake, fake, fake, 1, fake, 8,
                                         None) == _lib.ROCBTA_ERR_WORKSPACE


def test_ddbt_workspace_and_argument_errors(lib):
    assert lib.rocbta_workspace_bytes(_lib.OP_DDBTASC, 4, 256, 32, 0) >= 7 * 256 * 256 * 8
    assert lib.rocbta_workspace_bytes(_lib.OP_DDBTASCI, 4, 256, 32, 0) > 0
    fake = ctypes.c_void_p(16)
    big = 1 << 24
    # upward sweeps exist for the arrowless form only
    assert lib.rocbta_dddbtasc(3, 4, 2, fake, fake, fake, fake, fake, fake, 1, 1, fake, big,
                               None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci(3, 4, 0, fake, fake, fake, None, None, None, 1, 2, fake, big,
                                None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasc(3, 4, 2, fake, fake, None, fake, fake, fake, 1, 0, fake, big,
                               None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci(3, 4, 2, fake, fake, fake, fake, fake, fake, 1, 0, fake, 8,
                                None) == _lib.ROCBTA_ERR_WORKSPACE


def test_ddbt_permuted_argument_errors(lib):
    fake = ctypes.c_void_p(16)
    big = 1 << 24
    # n < 2, a missing buffer, then a workspace too small
    assert lib.rocbta_dddbtasc_permuted(1, 4, 2, *([fake] * 8), fake, big,
                                        None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci_permuted(3, 4, 2, *([fake] * 7), None, fake, big,
                                         None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasc_permuted(3, 4, 0, fake, fake, fake, None, None, None, fake, fake,
                                        fake, 8, None) == _lib.ROCBTA_ERR_WORKSPACE


def test_ddbt_quadratic_permuted_argument_errors(lib):
    fake = ctypes.c_void_p(16)
    big = 1 << 24
    # a missing B buffer, n < 2, then a workspace too small
    assert lib.rocbta_dddbtasc_quadratic_permuted(3, 4, 2, *([fake] * 15), None, fake, big,
                                                  None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci_quadratic_permuted(1, 4, 2, *([fake] * 16), fake, big,
                                                   None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci_quadratic_permuted(3, 4, 2, *([fake] * 16), fake, 8,
                                                   None) == _lib.ROCBTA_ERR_WORKSPACE


def test_ddbt_quadratic_argument_errors(lib):
    fake = ctypes.c_void_p(16)
    big = 1 << 26
    # missing B arrays, then a workspace too small
    assert lib.rocbta_dddbtasc_quadratic(3, 4, 2, fake, fake, fake, fake, fake, fake, None, fake,
                                         fake, fake, fake, fake, 1, 0, fake, big, None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci_quadratic(3, 4, 2, fake, fake, fake, fake, fake, fake, fake, fake,
                                          fake, fake, None, fake, 1, 0, fake, big, None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci_quadratic(3, 4, 2, *([fake] * 12), 1, 0, fake, 8,
                                          None) == _lib.ROCBTA_ERR_WORKSPACE
    # upward: the BT form only; an unknown direction
    assert lib.rocbta_dddbtasc_quadratic(3, 4, 2, *([fake] * 12), 1, 1, fake, big,
                                         None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci_quadratic(3, 4, 0, *([fake] * 12), 1, 2, fake, big,
                                          None) == _lib.ROCBTA_ERR_ARG
    assert lib.rocbta_dddbtasci_quadratic(3, 4, 0, *([fake] * 12), 1, 1, fake, 8,
                                          None) == _lib.ROCBTA_ERR_WORKSPACE
