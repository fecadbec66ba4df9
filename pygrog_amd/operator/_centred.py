"""Centred FFT plumbing without rolls (even grid sizes).

For an even length N (M = N/2): fftc(x)[k] = (-1)^(k-M) fftn((-1)^n x)[k]
and ifftc(x)[n] = (-1)^(n-M) ifftn((-1)^k x)[n] (the centred transforms of
src/pygrog/_base/_fftc.py:114-139; multiplying by +-1 is exact).  The
image-side signs are applied by the fused expand / combine kernels
(``pg_sense_expand_c64``, ``pg_sense_combine_c64``), the k-space-side signs
by the scatter / gather weights or the mask, so the operators call
torch.fft's plain transforms with no fftshift copies.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib


def even_grid(grid_shape) -> bool:
    return all(int(g) % 2 == 0 for g in grid_shape)


def half_sign(grid_shape) -> float:
    """prod over axes of (-1)^(N/2): the constant part of (-1)^(k - M)."""
    return float(np.prod([(-1.0) ** (int(g) // 2) for g in grid_shape]))


def cell_signs(flat_idx: torch.Tensor, grid_shape) -> torch.Tensor:
    """(-1)^(k0 + k1 + ...) of flat cell indices (float32, same shape)."""
    idx = flat_idx.to(torch.int64)
    par = torch.zeros_like(idx)
    rem = idx
    for g in reversed([int(v) for v in grid_shape]):
        par = par + rem % g
        rem = rem // g
    return (1 - 2 * (par & 1)).to(torch.float32)


def grid_signs(grid_shape, device) -> torch.Tensor:
    """(-1)^(k0 + k1 + ...) over the whole grid (float32, grid_shape)."""
    G = int(np.prod(grid_shape))
    return cell_signs(torch.arange(G, device=device), grid_shape).reshape(tuple(int(g) for g in grid_shape))


def expand(x: torch.Tensor, smaps: torch.Tensor | None, B: int, C: int, image_shape, grid_shape,
           modulate: bool = True) -> torch.Tensor:
    """(B, *image) with maps (C, *image), or (B, C, *image) without, ->
    (B, C, *grid): coil images zero-padded to the grid and sign-modulated."""
    dev = x.device
    out = torch.empty(B, C, *grid_shape, dtype=torch.complex64, device=dev)
    xs = x.to(torch.complex64).contiguous()
    sm = None if smaps is None else smaps.to(dev, torch.complex64).resolve_conj().contiguous()
    _lib.call("pg_sense_expand_c64", _lib.ptr(xs), _lib.ptr(sm), B, C, len(grid_shape), _lib.dims(image_shape),
              _lib.dims(grid_shape), int(modulate), _lib.ptr(out), _lib.stream(dev))
    return out


def combine(full: torch.Tensor, smaps: torch.Tensor | None, B: int, C: int, image_shape, grid_shape,
            sign: float = 1.0, modulate: bool = True) -> torch.Tensor:
    """(B, C, *grid) -> (B, *image) = sign * sum_c conj(S_c) crop(s * full_c),
    or without maps (B, C, *image) = sign * crop(s * full)."""
    dev = full.device
    f = full.to(torch.complex64).contiguous()
    if smaps is None:
        out = torch.empty(B, C, *image_shape, dtype=torch.complex64, device=dev)
        sm = None
    else:
        out = torch.empty(B, *image_shape, dtype=torch.complex64, device=dev)
        sm = smaps.to(dev, torch.complex64).resolve_conj().contiguous()
    _lib.call("pg_sense_combine_c64", _lib.ptr(f), _lib.ptr(sm), B, C, len(grid_shape), _lib.dims(image_shape),
              _lib.dims(grid_shape), int(modulate), float(sign), _lib.ptr(out), _lib.stream(dev))
    return out
