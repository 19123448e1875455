"""Utilities: synthetic matrices, conditioning, permutation buffers."""

from .permutation import allocate_pobtax_permutation_buffers, allocate_pobtx_permutation_buffers

__all__ = ["allocate_pobtax_permutation_buffers", "allocate_pobtx_permutation_buffers"]
