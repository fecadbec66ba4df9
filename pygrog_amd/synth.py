"""Seeded synthetic MRI workloads (trajectories, phantoms, coil maps,
calibration regions and FP64 non-uniform DFT k-space).

Used by the benchmark, the serving demo and the tests; nothing here is on
the gridding path.  Conventions:
  * coordinates in cycles/pixel in [-0.5, 0.5), float32, shape
    (*leading, ndim), axis order = grid axis order (axis 0 slowest);
  * images are on the centred grid r = -(n // 2) .. n - 1 - n // 2;
  * k-space of an image m:  y(k) = sum_r m(r) exp(-2 pi i k . r)  (FP64);
  * the calibration region is the centre of the centred orthonormal FFT of
    the coil images (the Cartesian samples the GRAPPA operators train on).
"""

from __future__ import annotations

import math

import numpy as np
import torch

GOLDEN_ANGLE_DEG = 111.246117975  # 180 / golden ratio
GOLDEN_MEANS = (0.4656, 0.6823)   # 3D golden means (Chan et al., MRM 61(2), 2009)


def radial_2d(n_spokes: int, n_read: int, offset: float = 0.0) -> np.ndarray:
    """Golden-angle radial, full-diameter spokes: (n_spokes, n_read, 2)."""
    r = (np.arange(n_read) - n_read // 2) / n_read
    th = np.deg2rad(GOLDEN_ANGLE_DEG) * np.arange(n_spokes) + offset
    return np.stack([np.outer(np.cos(th), r), np.outer(np.sin(th), r)], -1).astype(np.float32)


def kooshball_3d(n_spokes: int, n_read: int, first: int = 0) -> np.ndarray:
    """Golden-means 3D radial, full-diameter spokes: (n_spokes, n_read, 3).

    Spoke n points along (cos a sin p, sin a sin p, cos p) with
    cos p = frac(n phi1), a = 2 pi frac(n phi2).
    """
    n = np.arange(first, first + n_spokes, dtype=np.float64)
    cz = np.mod(n * GOLDEN_MEANS[0], 1.0)
    az = 2 * np.pi * np.mod(n * GOLDEN_MEANS[1], 1.0)
    sz = np.sqrt(1 - cz * cz)
    d = np.stack([cz, sz * np.sin(az), sz * np.cos(az)], -1)  # axis 0 = "z"
    r = (np.arange(n_read) - n_read // 2) / n_read
    return (d[:, None, :] * r[None, :, None]).astype(np.float32)


def spiral_vd_2d(n_interleaves: int, n_samples: int, fov: int, alpha: float = 2.0) -> np.ndarray:
    """Variable-density Archimedean spiral: radius k(t) = 0.5 t^alpha over
    t in [0, 1], enough turns to reach Nyquist at the edge for ``fov`` with
    ``n_interleaves`` arms; arm i rotated by 2 pi i / n_interleaves.
    Returns (n_interleaves, n_samples, 2) in cycles/pixel."""
    t = np.linspace(0, 1, n_samples, endpoint=False)
    turns = fov / (2 * n_interleaves)
    k = 0.5 * t**alpha
    th = 2 * np.pi * turns * t
    out = []
    for i in range(n_interleaves):
        ph = th + 2 * np.pi * i / n_interleaves
        out.append(np.stack([k * np.cos(ph), k * np.sin(ph)], -1))
    return np.asarray(out, dtype=np.float32)


def _grid_coords(shape, device):
    axes = [(torch.arange(s, device=device, dtype=torch.float32) - s // 2) / s for s in shape]
    return torch.stack(torch.meshgrid(*axes, indexing="ij"), -1)


def ellipse_phantom(shape, n: int = 10, seed: int = 0, device="cpu") -> torch.Tensor:
    """Sum of ``n`` random ellipses/ellipsoids, intensity U[0.1, 1], real
    float32 on ``device`` (parameters drawn on the host, seeded)."""
    rng = np.random.default_rng(seed)
    ndim = len(shape)
    x = _grid_coords(shape, device)
    img = torch.zeros(shape, dtype=torch.float32, device=device)
    for _ in range(n):
        c = torch.as_tensor(rng.uniform(-0.25, 0.25, ndim), dtype=torch.float32, device=device)
        ax = torch.as_tensor(rng.uniform(0.05, 0.3, ndim), dtype=torch.float32, device=device)
        if ndim == 2:
            a = rng.uniform(0, np.pi)
            R = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
        else:
            R, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        Rt = torch.as_tensor(R, dtype=torch.float32, device=device)
        u = (x - c) @ Rt
        img += float(rng.uniform(0.1, 1.0)) * (((u / ax) ** 2).sum(-1) <= 1.0).float()
    return img


def gaussian_coil_maps(shape, n_coils: int, sigma: float = 0.4, radius: float = 0.5, seed: int = 0,
                       device="cpu") -> torch.Tensor:
    """Gaussian-bump sensitivities (width ``sigma`` FOV) centred on a ring
    (2D) or a Fibonacci sphere (3D) of ``radius`` FOV, random phase per coil;
    unnormalised complex64 (n_coils, *shape) on ``device``."""
    rng = np.random.default_rng(seed)
    ndim = len(shape)
    x = _grid_coords(shape, device)
    maps = torch.empty((n_coils, *shape), dtype=torch.complex64, device=device)
    for c in range(n_coils):
        if ndim == 2:
            a = 2 * np.pi * c / n_coils
            ctr = radius * np.array([np.cos(a), np.sin(a)])
        else:
            z = 1 - 2 * (c + 0.5) / n_coils
            a = c * np.pi * (3 - np.sqrt(5))
            ctr = radius * np.array([z, np.sqrt(1 - z * z) * np.cos(a), np.sqrt(1 - z * z) * np.sin(a)])
        ct = torch.as_tensor(ctr, dtype=torch.float32, device=device)
        amp = torch.exp(-((x - ct) ** 2).sum(-1) / (2 * sigma**2))
        ph = float(rng.uniform(0, 2 * np.pi))
        maps[c] = amp * complex(np.cos(ph), np.sin(ph))
    return maps


def centered_fft(x: torch.Tensor, ndim: int) -> torch.Tensor:
    axes = tuple(range(-ndim, 0))
    return torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(x, dim=axes), dim=axes, norm="ortho"), dim=axes)


def acr(coil_images: torch.Tensor, size) -> torch.Tensor:
    """Centre ``size`` region of the centred FFT of (C, *shape) images."""
    ndim = coil_images.ndim - 1
    k = centered_fft(coil_images, ndim)
    sl = [slice(None)]
    for s, n in zip(coil_images.shape[1:], size):
        sl.append(slice(s // 2 - n // 2, s // 2 - n // 2 + n))
    return k[tuple(sl)]


def nudft(image: torch.Tensor, coords: np.ndarray, chunk: int = 8192) -> torch.Tensor:
    """FP64 non-uniform DFT of (C, *shape) images at coords (M, ndim):
    y_c(k) = sum_r m_c(r) exp(-2 pi i k . r), separable per axis, computed
    on the images' device in complex128 chunks of ``chunk`` samples."""
    dev = image.device
    img = image.to(torch.complex128)
    C = img.shape[0]
    shape = img.shape[1:]
    ndim = len(shape)
    k = torch.as_tensor(np.asarray(coords, np.float64).reshape(-1, ndim), device=dev)
    M = k.shape[0]
    r = [torch.arange(s, device=dev, dtype=torch.float64) - s // 2 for s in shape]
    out = torch.empty(C, M, dtype=torch.complex128, device=dev)
    for m0 in range(0, M, chunk):
        kk = k[m0:m0 + chunk]
        E = [torch.exp(-2j * math.pi * kk[:, d:d + 1] * r[d][None, :]) for d in range(ndim)]  # (m, n_d)
        # contract the last axis first: T = img @ E_last^T, then the others
        T = img.reshape(C, -1, shape[-1]) @ E[-1].T.contiguous()  # (C, prod(rest), m)
        if ndim == 2:
            T = T.reshape(C, shape[0], -1)  # (C, n0, m)
            out[:, m0:m0 + kk.shape[0]] = torch.einsum("cxm,mx->cm", T, E[0])
        else:
            T = T.reshape(C, shape[0], shape[1], -1)
            T = torch.einsum("cxym,my->cxm", T, E[1])
            out[:, m0:m0 + kk.shape[0]] = torch.einsum("cxm,mx->cm", T, E[0])
    return out


def complex_gaussian(shape, seed: int, device, dtype=torch.complex64) -> torch.Tensor:
    """i.i.d. complex Gaussian, real and imaginary parts each N(0, 1)."""
    g = torch.Generator(device=device).manual_seed(seed)
    re = torch.randn(shape, generator=g, device=device, dtype=torch.float32)
    im = torch.randn(shape, generator=g, device=device, dtype=torch.float32)
    return torch.complex(re, im).to(dtype)


def calibration(shape, n_coils: int, acr_size, seed: int, device, n_ellipses: int = 10) -> torch.Tensor:
    """ACR (C, *acr_size) complex64 from a seeded ellipse(oid) phantom times
    Gaussian coil maps (sigma 0.4 FOV)."""
    ph = ellipse_phantom(shape, n_ellipses, seed, device)
    maps = gaussian_coil_maps(shape, n_coils, 0.4, 0.5, seed + 1, device)
    return acr(maps * ph, acr_size).to(torch.complex64)
