"""numpy float32 restatement of the reference GROG plan (TEST INFRASTRUCTURE ONLY).

Every float expression keeps the reference's float32 operation order, each
op separately rounded (numpy ufuncs never fuse), so its target indices,
shifts and quantised table entries are the values the reference computes:

  rescale_coords      src/pygrog/calib/_grog.py:1318-1322
  kernel_offsets      src/pygrog/calib/_grog.py:1240-1277
  create_plan_single  src/pygrog/calib/_grog.py:1081-1234
  attach_fft_plan     src/pygrog/calib/_grog.py:918-998  (stable argsort: the
                      reference's unstable torch.argsort may order equal
                      indices differently; only tie order can differ)
  quantize            src/pygrog/calib/_grog.py:406-422 (pfac / stepsize /
                      nsteps from the precision or 1/q quantisation)

The exponential in the EGROG weights is numpy's float32 exp, so weights are
compared at a tolerance, not bitwise.
"""

from __future__ import annotations

import itertools

import numpy as np

f32 = np.float32


def kernel_offsets(ndim: int, width: int, shape: str) -> np.ndarray:
    lo = -((width - 1) // 2)
    axis = range(lo, lo + width)
    offs = np.array(list(itertools.product(axis, repeat=ndim)), dtype=np.int32).reshape(-1, ndim)
    s = shape.lower()
    if s in ("square", "cube"):
        return offs
    if s in ("circle", "sphere"):
        r = np.sqrt(np.sum(offs.astype(f32) ** 2, axis=-1))
        return offs[r <= (width - 1) / 2.0 + 0.5]
    if s == "cross":
        return offs[np.count_nonzero(offs, axis=-1) <= 1]
    raise ValueError(shape)


def default_oversamp(ndim, oversamp):
    if oversamp is None:
        return (1.0, 1.0) if ndim == 2 else (1.0, 1.0, 1.2)
    if np.isscalar(oversamp):
        return (float(oversamp),) * ndim
    return tuple(oversamp)


def rescale_coords(coords: np.ndarray, amp) -> np.ndarray:
    cmax = np.abs(coords).reshape(-1, coords.shape[-1]).max(axis=0)
    return 0.5 * np.asarray(amp, dtype=coords.dtype) * coords / cmax


def create_plan_single(shape, coords, oversamp=None, kernel_width=2, kernel_shape="circle") -> dict:
    coords = np.asarray(coords, dtype=f32)
    ndim = coords.shape[-1]
    shape = tuple(int(s) for s in (ndim * [shape] if np.isscalar(shape) else shape))
    oversamp = default_oversamp(ndim, oversamp)
    grid_shape = tuple(int(np.ceil(s * o)) for s, o in zip(shape, oversamp))
    amp = [2 * (s - 1 - s // 2) for s in shape]
    cs = rescale_coords(coords, amp)
    offsets = kernel_offsets(ndim, kernel_width, kernel_shape)
    origins = np.array([-(s // 2) for s in shape], dtype=f32)
    steps = np.array([(s - 1) / (gs - 1) if gs > 1 else 1.0 for s, gs in zip(shape, grid_shape)], dtype=f32)

    frac = (cs - origins) / steps
    nearest = np.rint(frac).astype(np.int32)
    tnd = nearest[..., None, :] + offsets  # (..., npts, kw, ndim)
    tcs = origins + tnd.astype(f32) * steps
    dist = tcs - cs[..., None, :]
    inb = np.all((tnd >= 0) & (tnd < np.array(grid_shape, dtype=np.int32)), axis=-1)
    inb &= np.all(np.abs(dist) <= 0.5, axis=-1)
    strides = np.array([int(np.prod(grid_shape[d + 1:])) for d in range(ndim)], dtype=np.int64)
    tflat = np.sum(tnd.astype(np.int64) * strides, axis=-1)
    tflat[~inb] = -1

    G = int(np.prod(grid_shape))
    dg = dist / steps
    dsq = dg[..., 0] * dg[..., 0]
    for d in range(1, ndim):
        dsq = dsq + dg[..., d] * dg[..., d]
    gauss = np.exp(f32(-0.5) * dsq).astype(f32)
    gauss[~inb] = 0.0
    flat_t = tflat.reshape(-1)
    flat_g = gauss.reshape(-1)
    valid = flat_t >= 0
    gsum = np.zeros(G, dtype=np.float64)
    np.add.at(gsum, flat_t[valid], flat_g[valid].astype(np.float64))
    w = np.zeros_like(flat_g)
    w[valid] = flat_g[valid] / np.maximum(gsum[flat_t[valid]], 1e-12).astype(f32)
    return {
        "shape": shape, "oversamp": oversamp, "grid_shape": grid_shape, "radius": 0.5,
        "target_idx": tflat, "distances": dist.astype(f32), "weights": w.reshape(tflat.shape),
        "offsets": offsets, "ndim": ndim,
    }


def attach_fft_plan(plan: dict) -> dict:
    idx = plan["target_idx"].reshape(-1).astype(np.int64).copy()
    w = plan["weights"].reshape(-1).astype(f32).copy()
    sqrt_w = np.sqrt(np.maximum(w, f32(0.0)))
    bad = idx < 0
    idx[bad] = 0
    sqrt_w[bad] = 0.0
    sort_perm = np.argsort(idx, kind="stable")
    inv_perm = np.empty_like(sort_perm)
    inv_perm[sort_perm] = np.arange(idx.size)
    plan = dict(plan)
    plan.update(indices=idx[sort_perm], sqrt_weights=sqrt_w[sort_perm], sort_perm=sort_perm, inv_perm=inv_perm,
                n_samples=int(idx.size))
    return plan


def quant_params(radius=0.5, precision=1, quant=None):
    """(pfac, stepsize, nsteps) as Python floats/ints (reference
    _grog.py:406-408, _grappa.py:100-101; ``quant`` = 1/q-grid steps)."""
    if quant is not None:
        return float(quant), 1.0 / quant, int(quant) + 1
    stepsize = 10 ** (-precision)
    return 10.0**precision, stepsize, int(2 * radius / stepsize + 1)


def quantize(distances: np.ndarray, n_entries: int, radius=0.5, precision=1, quant=None) -> np.ndarray:
    pfac, stepsize, nsteps = quant_params(radius, precision, quant)
    pf, st, r = f32(pfac), f32(stepsize), f32(radius)
    d = np.asarray(distances, dtype=f32)
    a = np.rint(d * pf) / pf
    q = np.rint((r + a) / st)
    strides = np.array([nsteps**k for k in range(d.shape[-1])], dtype=f32)
    e = np.rint(q * strides).astype(np.int32).astype(np.int64).sum(axis=-1)
    return np.clip(e, 0, n_entries - 1)


def density(plan: dict) -> np.ndarray:
    """sum of sqrt_w^2 per cell in sorted order, float32 (reference
    _grog.py:810-815)."""
    G = int(np.prod(plan["grid_shape"]))
    out = np.zeros(G, dtype=f32)
    sw = plan["sqrt_weights"]
    from .ops import psf_scatter_scalar

    psf_scatter_scalar(out, plan["indices"], (sw * sw).astype(f32))
    return out
