"""Triangular solve with the BT block Cholesky factor (drop-in for serinv.algs.pobts).

Reference: src/serinv/algs/pobts.py:10-54 (signature, kwargs, dispatch) and
:57-118 (algorithm); runs as rocbta_dpobtas with a = 0.
"""

from ._common import bta_shapes
from .pobtas import _solve


def pobts(L_diagonal_blocks, L_lower_diagonal_blocks, B, trans="N", **kwargs):
    """Solve with the BT block Cholesky factor, in place on B (shape (n*b, r) or (n*b,))."""
    D, L = L_diagonal_blocks, L_lower_diagonal_blocks
    n, b, a, dev = bta_shapes(D, L, None, None)
    _solve(D, L, None, None, B, trans, kwargs, n, b, 0, dev, "pobts")
