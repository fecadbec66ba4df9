"""Pins of the oracle's FFT-side operators (CPU only): closed forms of the
centred DFT, dense brute-force matrices on tiny grids, and CG against a
dense direct solve.  A wrong shift direction, exponent sign, pad offset or
permutation fails one of these."""

import numpy as np
import pytest

from oracle import operators as oop


def _centred_dft_matrix(M: int) -> np.ndarray:
    """F[k, j] = exp(-2 pi i (k - M//2)(j - M//2) / M) / sqrt(M) (even M)."""
    k = np.arange(M) - M // 2
    return np.exp(-2j * np.pi * np.outer(k, k) / M) / np.sqrt(M)


def test_fftc_delta_is_flat():
    for shape in [(8,), (6, 10), (4, 4, 6)]:
        x = np.zeros(shape, np.complex128)
        x[tuple(s // 2 for s in shape)] = 1.0
        X = oop.fftc(x, len(shape))
        np.testing.assert_allclose(X, np.full(shape, 1 / np.sqrt(np.prod(shape))), atol=1e-14)


@pytest.mark.parametrize("M", [6, 8, 12])
def test_fftc_matches_closed_form_and_inverse(M, rng):
    F = _centred_dft_matrix(M)
    x = rng.standard_normal(M) + 1j * rng.standard_normal(M)
    np.testing.assert_allclose(oop.fftc(x, 1), F @ x, atol=1e-12)
    np.testing.assert_allclose(oop.ifftc(x, 1), F.conj().T @ x, atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(oop.fftc(x, 1)), np.linalg.norm(x), rtol=1e-12)


def test_resize_crop_and_pad_offsets():
    x = np.arange(7.0)
    np.testing.assert_array_equal(oop.resize(x, (4,)), [1, 2, 3, 4])  # start (7-4)//2 = 1
    np.testing.assert_array_equal(oop.resize(np.arange(4.0), (7,)), [0, 0, 1, 2, 3, 0, 0])  # before = 3//2 = 1
    np.testing.assert_array_equal(oop.resize(np.arange(1.0, 4.0), (6,)), [0, 1, 2, 3, 0, 0])  # before = 3//2 = 1
    y = np.arange(30.0).reshape(5, 6)
    np.testing.assert_array_equal(oop.resize(y, (3, 2)), y[1:4, 2:4])


def _sparse_setup(rng, grid=(8, 8), image=(6, 6), n=40):
    G = int(np.prod(grid))
    idx = rng.integers(0, G, size=n)
    w = rng.uniform(0.5, 2.0, size=n)
    sqrt_w = np.sqrt(w)
    perm = np.argsort(idx, kind="stable")
    inv = np.empty_like(perm)
    inv[perm] = np.arange(n)
    return idx, sqrt_w, perm, inv


def _dense_sparse_matrix(idx, sqrt_w, grid, image):
    """A[s, pixel]: pad the image at offset (g-i)//2, centred DFT, pick cell idx[s]."""
    F = np.kron(_centred_dft_matrix(grid[0]), _centred_dft_matrix(grid[1]))
    P = np.zeros((grid[0] * grid[1], image[0] * image[1]))
    o0, o1 = (grid[0] - image[0]) // 2, (grid[1] - image[1]) // 2
    for a in range(image[0]):
        for b in range(image[1]):
            P[(a + o0) * grid[1] + (b + o1), a * image[1] + b] = 1.0
    return (sqrt_w[:, None] * (F @ P)[idx])


def test_sparse_forward_adjoint_match_dense_matrix(rng):
    grid, image = (8, 8), (6, 6)
    idx, sqrt_w, perm, inv = _sparse_setup(rng, grid, image)
    A = _dense_sparse_matrix(idx, sqrt_w, grid, image)
    img = rng.standard_normal((2, *image)) + 1j * rng.standard_normal((2, *image))
    y = oop.sparse_forward(img, idx[perm], sqrt_w[perm], inv, grid, image)
    np.testing.assert_allclose(y, (A @ img.reshape(2, -1).T).T, atol=1e-12)
    x = rng.standard_normal((2, idx.size)) + 1j * rng.standard_normal((2, idx.size))
    back = oop.sparse_adjoint(x, idx[perm], sqrt_w[perm], perm, grid, image)
    np.testing.assert_allclose(back.reshape(2, -1), (A.conj().T @ x.T).T, atol=1e-12)


def test_sparse_smaps_combine_is_adjoint(rng):
    grid, image = (8, 8), (6, 6)
    idx, sqrt_w, perm, inv = _sparse_setup(rng, grid, image)
    smaps = rng.standard_normal((3, *image)) + 1j * rng.standard_normal((3, *image))
    x = rng.standard_normal(image) + 1j * rng.standard_normal(image)
    y = rng.standard_normal((3, idx.size)) + 1j * rng.standard_normal((3, idx.size))
    Ax = oop.sparse_forward(x, idx[perm], sqrt_w[perm], inv, grid, image, smaps)
    AHy = oop.sparse_adjoint(y, idx[perm], sqrt_w[perm], perm, grid, image, smaps)
    assert AHy.shape == image
    np.testing.assert_allclose(np.vdot(Ax, y), np.vdot(x, AHy), rtol=1e-12)


def test_masked_operators_match_dense_and_adjoint(rng):
    grid, image = (8, 8), (6, 6)
    mask = (rng.random(grid) > 0.5).astype(np.float64)
    F = np.kron(_centred_dft_matrix(8), _centred_dft_matrix(8))
    img = rng.standard_normal((2, *image)) + 1j * rng.standard_normal((2, *image))
    k = oop.masked_forward(img, mask, image)
    pad = np.stack([oop.resize(i, grid) for i in img]).reshape(2, -1)
    np.testing.assert_allclose(k.reshape(2, -1), (mask.reshape(-1) * (F @ pad.T).T), atol=1e-12)
    y = rng.standard_normal((2, *grid)) + 1j * rng.standard_normal((2, *grid))
    np.testing.assert_allclose(np.vdot(k, y), np.vdot(img, oop.masked_adjoint(y, mask, image)), rtol=1e-12)


def test_cg_matches_direct_solve(rng):
    n = 24
    M = rng.standard_normal((40, n)) + 1j * rng.standard_normal((40, n))
    N = M.conj().T @ M
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    damp = 0.1
    x = oop.cg(lambda v: N @ v, b, damp=damp, max_iter=200, tol=1e-14)
    np.testing.assert_allclose(x, np.linalg.solve(N + damp * np.eye(n), b), rtol=1e-9, atol=1e-10)
