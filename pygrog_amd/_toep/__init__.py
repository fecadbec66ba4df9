"""Toeplitz normal operators (reference src/pygrog/_toep/__init__.py)."""

from ._grog_toep import GrogToeplitzOp

__all__ = ["GrogToeplitzOp"]
