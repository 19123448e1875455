"""Permutation buffers of the permuted ("middle partition") routines.

Counterparts of serinv.utils.allocate_pobtax_permutation_buffers
(src/serinv/utils/pobtax.py:10-29) and allocate_pobtx_permutation_buffers
(src/serinv/utils/pobtx.py:13-40): an uninitialised array like the diagonal
blocks, (n, b, b), whose block i the permuted routines use for the (0, i)
block of the factor or of the inverse.  For the non-SPD family,
allocate_ddbtx_permutation_buffers / allocate_ddbtax_permutation_buffers
(src/serinv/utils/ddbtx.py:9-37, ddbtax.py:8-13): zeroed buffers like the
lower blocks, for the fill-ins of row and column 0.
"""

import torch


def allocate_pobtx_permutation_buffers(A_diagonal_blocks, device_streaming=False):
    """Buffer for the permuted BT routines (pobtf / pobtsi / pobts with buffer=).

    device_streaming=True asks the reference for pinned host memory, for its
    streaming routines; those are not provided here, so it raises
    NotImplementedError.
    """
    if device_streaming:
        raise NotImplementedError("host-device streaming of the permuted forms is not provided")
    if not isinstance(A_diagonal_blocks, torch.Tensor):
        raise TypeError("A_diagonal_blocks must be a torch.Tensor on a ROCm device")
    return torch.empty_like(A_diagonal_blocks, memory_format=torch.contiguous_format)


def allocate_pobtax_permutation_buffers(A_diagonal_blocks, device_streaming=False):
    """Buffer for the permuted BTA routines (pobtaf / pobtasi / pobtas with
    buffer=); the same array as the BT one (pobtax.py:29)."""
    return allocate_pobtx_permutation_buffers(A_diagonal_blocks, device_streaming)


def allocate_ddbtx_permutation_buffers(A_lower_diagonal_blocks, quadratic=False):
    """Buffers of the non-SPD permuted routines (ddbtsc / ddbtsci with
    buffers=), src/serinv/utils/ddbtx.py:9-37: a dict of zero arrays like
    the lower blocks, (n-1, b, b), "A_lower_buffer_blocks" and
    "A_upper_buffer_blocks", plus the two "B_..." ones when quadratic."""
    if not isinstance(A_lower_diagonal_blocks, torch.Tensor):
        raise TypeError("A_lower_diagonal_blocks must be a torch.Tensor on a ROCm device")
    keys = ["A_lower_buffer_blocks", "A_upper_buffer_blocks"]
    if quadratic:
        keys += ["B_lower_buffer_blocks", "B_upper_buffer_blocks"]
    return {k: torch.zeros_like(A_lower_diagonal_blocks, memory_format=torch.contiguous_format)
            for k in keys}


def allocate_ddbtax_permutation_buffers(A_lower_diagonal_blocks, quadratic=False):
    """The BTA form's buffers (src/serinv/utils/ddbtax.py:8-13): the same dict."""
    return allocate_ddbtx_permutation_buffers(A_lower_diagonal_blocks, quadratic)
