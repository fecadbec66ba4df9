"""numpy restatement of the reference's FFT-side operators (TEST
INFRASTRUCTURE ONLY: imported by tests and never by the product).

Each function follows the reference's per-row loops, in float64/complex128
unless stated, with the scatter / gather primitives taken from the C oracle
(``grog_oracle.c``):

  fftc / ifftc        src/pygrog/_base/_fftc.py:114-139 (ifftshift -> fftn(ortho) -> fftshift)
  resize              src/pygrog/_utils.py:13-52 (centre crop / zero pad, floor((o-i)/2) before)
  sparse_adjoint      src/pygrog/operator/_sparse_fft.py:574-621 and :1154-1165 (scatter of the
                      sort-permuted row, IFFT at the grid, crop, optional sum_c conj(S_c) img_c)
  sparse_forward      src/pygrog/operator/_sparse_fft.py:703-789 and :1167-1177 (pad, FFT,
                      gather, unsort with inv_perm)
  masked_adjoint      src/pygrog/operator/_masked_fft.py:289-343 via _forward_single :345-376
                      (gridded k-space -> image: mask, IFFT, crop, SENSE combine)
  masked_forward      src/pygrog/operator/_masked_fft.py:382-430 via _adjoint_single :432-459
                      (image -> gridded k-space: SENSE expand, pad, FFT, mask)
  cg                  src/pygrog/_solve/_solvers.py:49-139, every scalar in float64

Pinned by: the closed form of the centred DFT on a delta (a flat spectrum
of height 1/sqrt(G)), Parseval, brute-force dense matrices on tiny grids and
CG against a dense direct solve (tests/test_oracle.py).
"""

from __future__ import annotations

import numpy as np

from . import ops


def _axes(ndim_total: int, ndim_fft: int):
    return tuple(range(ndim_total - ndim_fft, ndim_total))


def fftc(x: np.ndarray, ndim: int) -> np.ndarray:
    ax = _axes(x.ndim, ndim)
    return np.fft.fftshift(np.fft.fftn(np.fft.ifftshift(x, axes=ax), axes=ax, norm="ortho"), axes=ax)


def ifftc(x: np.ndarray, ndim: int) -> np.ndarray:
    ax = _axes(x.ndim, ndim)
    return np.fft.fftshift(np.fft.ifftn(np.fft.ifftshift(x, axes=ax), axes=ax, norm="ortho"), axes=ax)


def resize(x: np.ndarray, oshape) -> np.ndarray:
    """Centre crop / zero pad every axis to ``oshape``."""
    oshape = tuple(int(s) for s in oshape)
    if len(oshape) != x.ndim:
        raise ValueError("resize: rank mismatch")
    out = np.zeros(oshape, dtype=x.dtype)
    src, dst = [], []
    for i_len, o_len in zip(x.shape, oshape):
        if o_len <= i_len:
            s = (i_len - o_len) // 2
            src.append(slice(s, s + o_len))
            dst.append(slice(0, o_len))
        else:
            s = (o_len - i_len) // 2
            src.append(slice(0, i_len))
            dst.append(slice(s, s + i_len))
    out[tuple(dst)] = x[tuple(src)]
    return out


def sparse_adjoint(x, indices, sqrt_w, sort_perm, grid_shape, image_shape, smaps=None) -> np.ndarray:
    """(R, n) natural-order rows -> (R, *image) or, with smaps (R, *image),
    the coil-combined image.  ``indices`` / ``sqrt_w`` are in sorted order."""
    x = np.asarray(x, dtype=np.complex128)
    R, n = x.shape
    G = int(np.prod(grid_shape))
    imgs = []
    for r in range(R):
        grid = np.zeros(G, np.complex128)
        ops.scatter_add(grid, x[r][sort_perm], indices, np.asarray(sqrt_w, np.float64))
        imgs.append(resize(ifftc(grid.reshape(grid_shape), len(grid_shape)), image_shape))
    imgs = np.stack(imgs)
    if smaps is not None:
        return (imgs * np.conj(np.asarray(smaps, np.complex128))).sum(0)
    return imgs


def sparse_forward(img, indices, sqrt_w, inv_perm, grid_shape, image_shape, smaps=None) -> np.ndarray:
    """(R, *image) (or one image with smaps) -> (R, n) natural-order rows."""
    img = np.asarray(img, dtype=np.complex128)
    if smaps is not None:
        img = img[None] * np.asarray(smaps, np.complex128)
    rows = []
    for r in range(img.shape[0]):
        k = fftc(resize(img[r], grid_shape), len(grid_shape)).reshape(-1)
        rows.append(ops.gather(k, indices, np.asarray(sqrt_w, np.float64))[inv_perm])
    return np.stack(rows)


def masked_adjoint(kgrid, mask, image_shape, smaps=None) -> np.ndarray:
    """Gridded k-space (C, *grid) -> image (C, *image) or combined image."""
    k = np.asarray(kgrid, np.complex128) * np.asarray(mask, np.float64)[None]
    nd = mask.ndim
    imgs = np.stack([resize(ifftc(k[c], nd), image_shape) for c in range(k.shape[0])])
    if smaps is not None:
        return (imgs * np.conj(np.asarray(smaps, np.complex128))).sum(0)
    return imgs


def masked_forward(img, mask, image_shape, smaps=None) -> np.ndarray:
    """Image (C, *image) (or one image with smaps) -> masked k-space grid."""
    img = np.asarray(img, dtype=np.complex128)
    if smaps is not None:
        img = img[None] * np.asarray(smaps, np.complex128)
    grid_shape = mask.shape
    return np.stack([fftc(resize(img[c], grid_shape), mask.ndim) * mask for c in range(img.shape[0])])


def cg(normal, rhs, *, damp=0.0, max_iter=20, tol=1e-6) -> np.ndarray:
    """Conjugate gradients on ``(N + damp I) x = rhs`` with float64 scalars;
    ``normal`` maps an array to an array of the same shape."""
    x = np.zeros_like(rhs, dtype=np.complex128)
    r = rhs.astype(np.complex128) - normal(x) - damp * x
    p = r.copy()
    rz_old = float(np.vdot(r, r).real)
    r0 = float(np.linalg.norm(r)) + 1e-30
    for _ in range(max_iter):
        Ap = normal(p) + damp * p
        alpha = rz_old / (float(np.vdot(p, Ap).real) + 1e-30)
        x = x + alpha * p
        r = r - alpha * Ap
        if float(np.linalg.norm(r)) / r0 < tol:
            break
        rz_new = float(np.vdot(r, r).real)
        p = r + (rz_new / (rz_old + 1e-30)) * p
        rz_old = rz_new
    return x
