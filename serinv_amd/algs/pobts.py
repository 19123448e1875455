"""Triangular solve with the BT block Cholesky factor (drop-in for serinv.algs.pobts).

Reference: src/serinv/algs/pobts.py:10-54 (signature, kwargs, dispatch),
:57-118 (algorithm) and :121-165 (permuted algorithm, with buffer=); runs as
rocbta_dpobtas / rocbta_dpobtas_permuted with a = 0.
"""

from ._common import bta_shapes
from .pobtas import _solve


def pobts(L_diagonal_blocks, L_lower_diagonal_blocks, B, trans="N", **kwargs):
    """Solve with the BT block Cholesky factor, in place on B (shape (n*b, r) or (n*b,))."""
    D, L = L_diagonal_blocks, L_lower_diagonal_blocks
    n, b, a, dev = bta_shapes(D, L, None, None)
    _solve(D, L, None, None, B, trans, kwargs, n, b, 0, dev, "pobts")
