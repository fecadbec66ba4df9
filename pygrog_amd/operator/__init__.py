"""Sparse and masked FFT operators (reference src/pygrog/operator/__init__.py)."""

from ._masked_fft import MaskedFFT, MaskedFFTPlan
from ._sparse_fft import SparseFFT, gather, scatter_add

__all__ = ["MaskedFFT", "MaskedFFTPlan", "SparseFFT", "gather", "scatter_add"]
