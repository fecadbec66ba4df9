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


def sparse_reference(y: np.ndarray, interp_idx: np.ndarray, table: np.ndarray, kw: int,
                     dtype=np.complex64) -> np.ndarray:
    """y (C, N) -> sparse (C, N*kw): every replica (n, k) gets K[interp_idx] @ y[:, n]
    (complex64, or complex128 with ``dtype``)."""
    C, N = y.shape
    rep = np.repeat(y.T[:, None, :], kw, axis=1).reshape(N * kw, 1, C)
    out = ops.interpolate(rep.astype(dtype), interp_idx.reshape(-1), table.astype(dtype))
    return np.ascontiguousarray(out[:, 0, :].T)


def grid_reference(y, plan: dict, interp_idx, table, cells: np.ndarray | None = None, dtype=np.complex64):
    """-> (grid (C, G), density (G,) float32), the grid in ``dtype``
    (complex64, or complex128: the same flow with float64 weights, as the
    reference's ops dispatch on the data type).

    With ``cells`` (complex64 only): only those cells, -> (C, len(cells))
    complex64 (columns in the given order).  Only the replicas landing there
    are formed, with the same per-replica arithmetic, and each cell is
    summed over its replicas in the same (sorted) order as the full flow, so
    the columns equal the full grid's bit for bit."""
    if cells is not None:
        cells = np.asarray(cells, dtype=np.int64)
        if np.unique(cells).size != cells.size:
            raise ValueError("cells must be distinct")
        return _grid_reference_cells(y, plan, interp_idx, table, cells)
    kw = plan["offsets"].shape[0]
    rt = np.float32 if dtype == np.complex64 else np.float64
    sparse = sparse_reference(np.asarray(y, dtype=dtype), interp_idx, table, kw, dtype)
    pre_w = plan["sqrt_weights"][plan["inv_perm"]].astype(dtype)
    sparse_dc = (sparse * pre_w[None, :]).astype(dtype)
    G = int(np.prod(plan["grid_shape"]))
    C = sparse.shape[0]
    grid = np.zeros((C, G), dtype=dtype)
    for c in range(C):
        g = np.zeros(G, dtype=dtype)
        ops.scatter_add(g, sparse_dc[c][plan["sort_perm"]], plan["indices"], plan["sqrt_weights"].astype(rt))
        grid[c] = g
    from .plan import density

    return grid, density(plan)


def _grid_reference_cells(y, plan: dict, interp_idx, table, cells: np.ndarray) -> np.ndarray:
    kw = plan["offsets"].shape[0]
    G = int(np.prod(plan["grid_shape"]))
    pos = np.full(G, -1, dtype=np.int64)
    pos[cells] = np.arange(len(cells))
    ind = plan["indices"].reshape(-1)
    q = np.flatnonzero(pos[ind] >= 0)  # sorted positions, ascending = the full flow's order
    r = plan["sort_perm"].reshape(-1)[q]  # replica (n * kw + k)
    e = np.asarray(interp_idx).reshape(-1)[r]
    y = np.asarray(y)
    rows = np.ascontiguousarray(y[:, r // kw].T.astype(np.complex64))[:, None, :]
    applied = ops.interpolate(rows, e, np.asarray(table, dtype=np.complex64))[:, 0, :]  # (n, C)
    sw = plan["sqrt_weights"].reshape(-1)[q]
    # pre_weights[r] = sqrt_weights[inv_perm[r]] = sqrt_weights[q]
    data = (applied * sw.astype(np.complex64)[:, None]).astype(np.complex64)
    C = y.shape[0]
    out = np.zeros((C, len(cells)), dtype=np.complex64)
    col = pos[ind[q]]
    for c in range(C):
        g = np.zeros(len(cells), dtype=np.complex64)
        ops.scatter_add(g, data[:, c], col, sw)
        out[c] = g
    return out


def grid_fp64(y, plan: dict, interp_idx, table64, cells: np.ndarray | None = None,
              chunk: int = 4096) -> np.ndarray:
    """FP64 oracle grid (C, G); with ``cells`` only those cells are computed
    (columns in the given order, shape (C, len(cells))).  ``y`` may be
    complex64; it is widened per replica."""
    y = np.asarray(y)
    C = y.shape[0]
    kw = plan["offsets"].shape[0]
    tflat = plan["target_idx"].reshape(-1)
    w = plan["weights"].reshape(-1).astype(np.float64)
    e = np.asarray(interp_idx).reshape(-1)
    valid = np.nonzero(tflat >= 0)[0]
    G = int(np.prod(plan["grid_shape"]))
    if cells is not None:
        if np.unique(cells).size != np.asarray(cells).size:
            raise ValueError("cells must be distinct")
        pos = np.full(G, -1, dtype=np.int64)
        pos[np.asarray(cells)] = np.arange(len(cells))
        valid = valid[pos[tflat[valid]] >= 0]
        col = pos[tflat[valid]]
        ncol = len(cells)
    else:
        col = tflat[valid]
        ncol = G
    src = valid // kw
    ev = e[valid]
    wv = w[valid]
    T = table64
    # V[r] = w_r K64[e_r] y[:, src_r], in chunks of replicas (only the
    # selected source columns are widened to complex128)
    V = np.empty((len(valid), C), dtype=np.complex128)
    for r0 in range(0, len(valid), chunk):
        r1 = min(r0 + chunk, len(valid))
        Tk = np.asarray(T[ev[r0:r1]], dtype=np.complex128)
        yk = np.asarray(y[:, src[r0:r1]], dtype=np.complex128).T
        V[r0:r1] = np.einsum("rij,rj->ri", Tk, yk) * wv[r0:r1, None]
    out = np.zeros((C, ncol), dtype=np.complex128)
    for c in range(C):
        out[c] = np.bincount(col, weights=V[:, c].real, minlength=ncol) + 1j * np.bincount(
            col, weights=V[:, c].imag, minlength=ncol)
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
