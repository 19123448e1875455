"""Permutation buffers of the permuted ("middle partition") routines.

Counterparts of serinv.utils.allocate_pobtax_permutation_buffers
(src/serinv/utils/pobtax.py:10-29) and allocate_pobtx_permutation_buffers
(src/serinv/utils/pobtx.py:13-40): an uninitialised array like the diagonal
blocks, (n, b, b), whose block i the permuted routines use for the (0, i)
block of the factor or of the inverse.
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
