"""Centered, orthonormal FFT / IFFT on torch.fft (reference
src/pygrog/_base/_fftc.py:17-139): ``fftshift(fftn(ifftshift(resize(x))))``."""

from __future__ import annotations

__all__ = ["fft", "ifft"]

import numpy as np
import torch

from .._utils import normalize_axes, resize


def _centered(x, oshape, axes, norm, inverse: bool):
    axes = normalize_axes(x.ndim, axes)
    tmp = resize(x, x.shape if oshape is None else oshape)
    tmp = torch.fft.ifftshift(tmp, dim=axes)
    tmp = (torch.fft.ifftn if inverse else torch.fft.fftn)(tmp, dim=axes, norm=norm)
    return torch.fft.fftshift(tmp, dim=axes)


def _prep(input, axes):
    if axes is not None and np.isscalar(axes):
        axes = (axes,)
    is_np = isinstance(input, np.ndarray)
    x = torch.as_tensor(input)
    if not torch.is_complex(x):
        x = x.to(torch.complex64)
    return x, axes, is_np


def fft(input, oshape=None, axes=None, center: bool = True, norm: str | None = "ortho"):  # noqa: A002
    """Centered FFT over ``axes`` (all axes if None)."""
    x, axes, is_np = _prep(input, axes)
    if center:
        out = _centered(x, oshape, axes, norm, inverse=False)
    else:
        out = torch.fft.fftn(x, s=oshape, dim=axes, norm=norm)
    if out.dtype != x.dtype:
        out = out.to(x.dtype)
    return out.numpy() if is_np else out


def ifft(input, oshape=None, axes=None, center: bool = True, norm: str | None = "ortho"):  # noqa: A002
    """Centered inverse FFT over ``axes`` (all axes if None)."""
    x, axes, is_np = _prep(input, axes)
    if center:
        out = _centered(x, oshape, axes, norm, inverse=True)
    else:
        out = torch.fft.ifftn(x, s=oshape, dim=axes, norm=norm)
    if out.dtype != x.dtype:
        out = out.to(x.dtype)
    return out.numpy() if is_np else out
