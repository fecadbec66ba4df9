"""Centered FFT helpers (FFT itself is torch.fft)."""

from ._fftc import fft, ifft  # noqa: F401

__all__ = ["fft", "ifft"]
