"""serinv-compatible sequential SPD BTA / BT routines on MI355X.

Drop-in counterparts of serinv.algs (src/serinv/algs/__init__.py:5-31) for
the SPD path and the diagonally dominant ddbt* family: same names, argument
order, in-place block layout, keyword arguments and error types, on float64
torch tensors on a ROCm device.
"""

from ._common import check_status, config
from .ddbtasc import ddbtasc, ddbtasci, ddbtsc, ddbtsci
from .pobtaf import pobtaf
from .pobtas import pobtas
from .pobtasi import pobtasi
from .pobtf import pobtf
from .pobts import pobts
from .pobtsi import pobtsi

__all__ = ["pobtaf", "pobtasi", "pobtas", "pobtf", "pobtsi", "pobts", "ddbtasc", "ddbtasci",
           "ddbtsc", "ddbtsci", "config", "check_status"]
