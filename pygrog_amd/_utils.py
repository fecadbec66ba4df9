"""Array helpers: centre crop / zero pad, axis normalisation, coordinate
rescaling (reference src/pygrog/_utils.py:13-137)."""

from __future__ import annotations

__all__ = ["estimate_shape", "normalize_axes", "rescale_coords", "resize"]

import math

import numpy as np
import torch


def resize(input: torch.Tensor, oshape) -> torch.Tensor:  # noqa: A002
    """Centre crop / zero pad every axis to ``oshape`` (reference
    _utils.py:13-52: crop start ``(i - o) // 2``, pad before ``(o - i) // 2``)."""
    ishape = tuple(input.shape)
    oshape = tuple(int(s) for s in oshape)
    if len(ishape) != len(oshape):
        raise ValueError(f"Input ndim ({len(ishape)}) != output ndim ({len(oshape)})")
    out = input
    pads = []
    for axis, (i_len, o_len) in enumerate(zip(ishape, oshape)):
        if o_len < i_len:
            out = out.narrow(axis, (i_len - o_len) // 2, o_len)
    # one pad call for all growing axes (torch pads from the last axis)
    need_pad = any(o > i for i, o in zip(ishape, oshape))
    if need_pad:
        for axis in range(len(oshape) - 1, -1, -1):
            i_len, o_len = out.shape[axis], oshape[axis]
            before = (o_len - i_len) // 2 if o_len > i_len else 0
            after = (o_len - i_len - before) if o_len > i_len else 0
            pads += [before, after]
        out = torch.nn.functional.pad(out, pads)
    return out


def normalize_axes(ndim: int, axes) -> list[int]:
    if axes is None:
        return list(range(ndim))
    if np.isscalar(axes):
        axes = (axes,)
    return [a % ndim for a in axes]


def rescale_coords(coords, amp):
    """Scale coordinates to ``(-0.5 amp, 0.5 amp)`` by the per-row max
    (reference _utils.py:79-112)."""
    if isinstance(coords, torch.Tensor):
        cmax = coords.abs().reshape(coords.shape[0], -1).max(dim=1).values
        amp_t = (
            torch.full((coords.shape[0],), amp, dtype=coords.dtype, device=coords.device)
            if np.isscalar(amp)
            else torch.as_tensor(amp, dtype=coords.dtype, device=coords.device)
        )
        return 0.5 * amp_t[:, None] * coords / cmax[:, None]
    cmax = np.abs(coords).reshape(coords.shape[0], -1).max(axis=1)
    amp = np.asarray([amp] * coords.shape[0] if np.isscalar(amp) else amp, dtype=coords.dtype)
    return 0.5 * amp[:, None] * coords / cmax[:, None]


def estimate_shape(coords) -> tuple[int, ...]:
    """Grid shape ``2 ceil(max |coord|)`` per axis (reference _utils.py:115-137)."""
    ndim = coords.shape[0]
    shape = []
    for i in range(ndim):
        c = np.asarray(coords[i]).ravel()
        n = max(int(2 * math.ceil(float(np.max(np.abs(c))))), 1)
        shape.append(n)
    return tuple(shape)
