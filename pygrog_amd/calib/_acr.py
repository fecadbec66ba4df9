"""Calibration-region extraction (drop-in for the reference's
src/pygrog/calib/_acr.py:12-113).

Index selection on tensors (any device; the data stay where they are):
Cartesian data are centre-cropped to ``cal_width`` per spatial axis;
non-Cartesian data keep the readout positions whose samples lie within
``cal_width / 2`` grid units of the k-space centre on every view (after
rescaling the trajectory to +-shape/2 by its global maximum), and for a
stack of stars (integer last coordinate) also the partitions within
``cal_width / 2``.
"""

from __future__ import annotations

__all__ = ["extract_acr"]

import numpy as np
import torch


def _center_crop(arr, oshape):
    """Centre crop the trailing axes to ``oshape`` (reference :116-129)."""
    oshape = list(oshape)
    if len(oshape) < arr.ndim:
        oshape = list(arr.shape[: arr.ndim - len(oshape)]) + oshape
    slices = []
    for i_len, o_len in zip(arr.shape, oshape):
        if o_len <= i_len:
            start = (i_len - o_len) // 2
            slices.append(slice(start, start + o_len))
        else:
            slices.append(slice(None))
    return arr[tuple(slices)]


def _rescale(coords: torch.Tensor, amp) -> torch.Tensor:
    """Scale to +-amp/2 by the maximum over each leading-axis slice
    (reference src/pygrog/_utils.py:79-112)."""
    cmax = coords.abs().reshape(coords.shape[0], -1).max(dim=1).values
    if np.isscalar(amp):
        amp_t = torch.full((coords.shape[0],), float(amp), dtype=coords.dtype, device=coords.device)
    else:
        amp_t = torch.as_tensor(amp, dtype=coords.dtype, device=coords.device)
    shape = (-1,) + (1,) * (coords.ndim - 1)
    return 0.5 * amp_t.reshape(shape) * coords / cmax.reshape(shape)


def extract_acr(data, cal_width: int = 20, coords=None, weights=None, shape=None, ndim=None, mask=None):
    """Calibration region of ``data`` (reference signature and returns):
    Cartesian -> ``cal_data`` or ``(cal_data, cal_mask)``; non-Cartesian ->
    ``(cal_data, cal_coords, cal_weights)``."""
    is_np = isinstance(data, np.ndarray)
    d = torch.as_tensor(data)
    if coords is None:
        if ndim is None:
            raise ValueError("Please provide number of spatial dimensions for Cartesian datasets")
        cal_shape = list(d.shape[:-ndim]) + ndim * [cal_width]
        out = _center_crop(d, cal_shape)
        if mask is not None:
            m = _center_crop(torch.as_tensor(mask), ndim * [cal_width])
            return (out.numpy(), m.numpy()) if is_np else (out, m)
        return out.numpy() if is_np else out
    if shape is None:
        raise ValueError("Please provide matrix size for Non Cartesian datasets")
    c = _rescale(torch.as_tensor(coords).to(d.device), shape)
    w = None if weights is None else torch.as_tensor(weights).to(d.device)
    if torch.allclose(c[..., -1], torch.round(c[..., -1])):
        # stack of stars: in-plane radius per readout position, partitions by |kz|
        cal_idx = (c[..., :2] ** 2).sum(dim=-1) ** 0.5 <= (0.5 * cal_width)
        cal_idx = cal_idx.reshape(-1, cal_idx.shape[-1]).prod(dim=0).bool()
        cal_z = c[..., -1].abs() <= (0.5 * cal_width)
        cal_z = cal_z.reshape(-1, cal_idx.shape[-1]).prod(dim=-1).bool()
        out_d = d[..., cal_idx][..., cal_z, :]
        out_c = c[..., cal_idx, :][..., cal_z, :, :]
        out_w = None if w is None else w[..., cal_idx][..., cal_z, :]
    else:
        cal_idx = (c ** 2).sum(dim=-1) ** 0.5 <= (0.5 * cal_width)
        cal_idx = cal_idx.reshape(-1, cal_idx.shape[-1]).prod(dim=0).bool()
        out_d = d[..., cal_idx]
        out_c = c[..., cal_idx, :]
        out_w = None if w is None else w[..., cal_idx]
    if is_np:
        return out_d.numpy(), out_c.numpy(), None if out_w is None else out_w.numpy()
    return out_d, out_c, out_w
