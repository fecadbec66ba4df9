"""Gridding oracles (TEST INFRASTRUCTURE ONLY).

``grid_reference``: the reference's ``interpolate(grid=True)`` flow
(src/pygrog/calib/_grog.py:437-550 -> _sparse_full_single :668-695 ->
_grid_kspace :736-902) on the C restatement of its CPU ops: every kw replica
gets its operator (``grog_interpolate``), the sparse data is multiplied by
``pre_weights`` (complex64 product), put in sorted order and scatter-added
with ``sqrt_weights``; density = index_add of ``sqrt_weights**2``.

``grid_fp64``: the same sum in float64 with an FP64 table,
``grid[c, cell] = sum_r w_r (K64[e_r] y[:, src_r])[c]`` with ``w_r`` the plan
weight, the target of the FP32 / BF16 error figures.

``image_rss``: ``interpolate(ret_image=True)`` (_grog.py:518-534): centred
IFFT at the grid shape, image-space crop, root-sum-of-squares over coils.
"""

from __future__ import annotations

import numpy as np

from . import ops


def sparse_reference(y: np.ndarray, interp_idx: np.ndarray, table: np.ndarray, kw: int) -> np.ndarray:
    """y (C, N) -> sparse (C, N*kw): every replica (n, k) gets K[interp_idx] @ y[:, n]."""
    C, N = y.shape
    rep = np.repeat(y.T[:, None, :], kw, axis=1).reshape(N * kw, 1, C)
    out = ops.interpolate(rep.astype(np.complex64), interp_idx.reshape(-1), table.astype(np.complex64))
    return np.ascontiguousarray(out[:, 0, :].T)


def grid_reference(y, plan: dict, interp_idx, table) -> tuple[np.ndarray, np.ndarray]:
    """-> (grid (C, G) complex64, density (G,) float32)."""
    kw = plan["offsets"].shape[0]
    sparse = sparse_reference(np.asarray(y, dtype=np.complex64), interp_idx, table, kw)
    pre_w = plan["sqrt_weights"][plan["inv_perm"]].astype(np.complex64)
    sparse_dc = (sparse * pre_w[None, :]).astype(np.complex64)
    G = int(np.prod(plan["grid_shape"]))
    C = sparse.shape[0]
    grid = np.zeros((C, G), dtype=np.complex64)
    for c in range(C):
        g = np.zeros(G, dtype=np.complex64)
        ops.scatter_add(g, sparse_dc[c][plan["sort_perm"]], plan["indices"], plan["sqrt_weights"])
        grid[c] = g
    from .plan import density

    return grid, density(plan)


def grid_fp64(y, plan: dict, interp_idx, table64, cells: np.ndarray | None = None) -> np.ndarray:
    """FP64 oracle grid (C, G); with ``cells`` only those cells are computed
    (columns in the given order, shape (C, len(cells)))."""
    y = np.asarray(y, dtype=np.complex128)
    C = y.shape[0]
    kw = plan["offsets"].shape[0]
    tflat = plan["target_idx"].reshape(-1)
    w = plan["weights"].reshape(-1).astype(np.float64)
    e = np.asarray(interp_idx).reshape(-1)
    valid = np.nonzero(tflat >= 0)[0]
    G = int(np.prod(plan["grid_shape"]))
    if cells is not None:
        pos = np.full(G, -1, dtype=np.int64)
        pos[np.asarray(cells)] = np.arange(len(cells))
        valid = valid[pos[tflat[valid]] >= 0]
        col = pos[tflat[valid]]
        ncol = len(cells)
    else:
        col = tflat[valid]
        ncol = G
    src = valid // kw
    out = np.zeros((C, ncol), dtype=np.complex128)
    T = np.asarray(table64, dtype=np.complex128)
    order = np.argsort(e[valid], kind="stable")
    ev = e[valid][order]
    bounds = np.flatnonzero(np.diff(ev)) + 1
    starts = np.concatenate([[0], bounds])
    ends = np.concatenate([bounds, [len(ev)]])
    for s0, s1 in zip(starts, ends):
        sel = order[s0:s1]
        V = (T[ev[s0]] @ y[:, src[sel]]) * w[valid][sel][None, :]
        for c in range(C):
            out[c] += np.bincount(col[sel], weights=V[c].real, minlength=ncol) + 1j * np.bincount(
                col[sel], weights=V[c].imag, minlength=ncol)
    return out


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def image_rss(grid: np.ndarray, grid_shape, image_shape) -> np.ndarray:
    """RSS image of a (C, G) grid (centred orthonormal IFFT, centre crop)."""
    C = grid.shape[0]
    g = grid.reshape(C, *grid_shape)
    axes = tuple(range(1, 1 + len(grid_shape)))
    img = np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(g, axes=axes), axes=axes, norm="ortho"), axes=axes)
    sl = tuple(slice((gs - s) // 2, (gs - s) // 2 + s) for gs, s in zip(grid_shape, image_shape))
    img = img[(slice(None),) + sl]
    return np.sqrt(np.sum(np.abs(img) ** 2, axis=0))
