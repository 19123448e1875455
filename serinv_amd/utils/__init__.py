"""Utilities: permutation buffers of the permuted routines."""

from .permutation import (allocate_ddbtax_permutation_buffers, allocate_ddbtx_permutation_buffers,
                          allocate_pobtax_permutation_buffers, allocate_pobtx_permutation_buffers)

__all__ = ["allocate_ddbtax_permutation_buffers", "allocate_ddbtx_permutation_buffers",
           "allocate_pobtax_permutation_buffers", "allocate_pobtx_permutation_buffers"]
