"""numpy front end of ``grog_oracle.c`` (TEST INFRASTRUCTURE ONLY).

Each function restates one reference CPU kernel; see the citations in
``grog_oracle.c``.  Arrays are converted to contiguous complex64/128 (or
float32/64) copies; in-place ops write back into the given array.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import lib

_P = ctypes.c_void_p
_I = ctypes.c_int64


def _sfx(dtype) -> tuple[str, type, type]:
    if np.dtype(dtype) in (np.complex64, np.float32):
        return "f32", np.complex64, np.float32
    return "f64", np.complex128, np.float64


def _p(a: np.ndarray):
    return a.ctypes.data_as(_P)


def interpolate(data: np.ndarray, indexes: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """out[n, b, :] = kernel[indexes[n]] @ data[n, b, :]; data (N, B, C)."""
    sfx, ct, _ = _sfx(data.dtype)
    d = np.ascontiguousarray(data, dtype=ct)
    k = np.ascontiguousarray(kernel, dtype=ct)
    idx = np.ascontiguousarray(indexes, dtype=np.int64)
    out = np.empty_like(d)
    N, B, C = d.shape
    getattr(lib(), f"or_interp_{sfx}")(_p(out), _p(d), _p(idx), _p(k), _I(N), _I(B), _I(C), _I(k.shape[0]))
    return out


def scatter_add(grid: np.ndarray, data: np.ndarray, indices: np.ndarray, weights: np.ndarray) -> None:
    sfx, ct, rt = _sfx(grid.dtype)
    assert grid.flags.c_contiguous and grid.dtype == ct
    d = np.ascontiguousarray(data, dtype=ct)
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    w = np.ascontiguousarray(weights, dtype=rt)
    getattr(lib(), f"or_scatter_add_{sfx}")(_p(grid), _I(grid.size), _p(d), _p(idx), _p(w), _I(idx.size))


def gather(grid: np.ndarray, indices: np.ndarray, weights: np.ndarray) -> np.ndarray:
    sfx, ct, rt = _sfx(grid.dtype)
    g = np.ascontiguousarray(grid, dtype=ct)
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    w = np.ascontiguousarray(weights, dtype=rt)
    out = np.empty(idx.size, dtype=ct)
    getattr(lib(), f"or_gather_{sfx}")(_p(out), _p(g), _I(g.size), _p(idx), _p(w), _I(idx.size))
    return out


def psf_scatter_scalar(psf: np.ndarray, indices: np.ndarray, w_sq: np.ndarray) -> None:
    sfx = "f32" if psf.dtype == np.float32 else "f64"
    assert psf.flags.c_contiguous
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    w = np.ascontiguousarray(w_sq, dtype=psf.dtype)
    getattr(lib(), f"or_psf_scalar_{sfx}")(_p(psf), _I(psf.size), _p(idx), _p(w), _I(idx.size))


def psf_scatter_outer(psf: np.ndarray, indices: np.ndarray, w_sq: np.ndarray, basis: np.ndarray) -> None:
    sfx, ct, rt = _sfx(psf.dtype)
    assert psf.flags.c_contiguous and psf.dtype == ct
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    w = np.ascontiguousarray(w_sq, dtype=rt)
    b = np.ascontiguousarray(basis, dtype=ct)
    getattr(lib(), f"or_psf_outer_{sfx}")(_p(psf), _I(psf.shape[0]), _I(psf.shape[1]), _p(idx), _p(w), _p(b),
                                           _I(idx.size))


def psf_scatter_outer_basis(psf, indices, w_sq, basis, time_index) -> None:
    sfx, ct, rt = _sfx(psf.dtype)
    assert psf.flags.c_contiguous and psf.dtype == ct
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    w = np.ascontiguousarray(w_sq, dtype=rt)
    b = np.ascontiguousarray(basis, dtype=ct)
    t = np.ascontiguousarray(time_index, dtype=np.int64)
    getattr(lib(), f"or_psf_outer_basis_{sfx}")(_p(psf), _I(psf.shape[0]), _I(psf.shape[1]), _p(idx), _p(w), _p(b),
                                                 _I(b.shape[1]), _p(t), _I(idx.size))
