"""Masked FFT operator on a gridded (Cartesian) k-space (drop-in for the
reference's src/pygrog/operator/_masked_fft.py).

``adjoint``: gridded k-space -> image (mask, centred IFFT at the grid shape,
image-space centre crop, optional SENSE combine); ``forward``: image ->
gridded k-space (zero pad, centred FFT, mask).  The reference loops over
coils and stack elements (:288-459); here each call is one batched pass on
the GPU.  ``MaskedFFTPlan`` is what ``GrogInterpolator.interpolate(grid=True)``
returns (mask, density = sum of squared GROG weights per cell).
"""

from __future__ import annotations

__all__ = ["MaskedFFT", "MaskedFFTPlan"]

import math

import numpy as np
import torch

from .. import _lib
from .._base._fftc import fft, ifft
from .._solve._mixin import SolveMixin
from .._utils import resize
from ._centred import combine, even_grid, expand, grid_signs, half_sign
from ._sparse_fft import numpy_io


class MaskedFFTPlan:
    """Plan of :class:`MaskedFFT` (reference _masked_fft.py:54-119)."""

    def __init__(self, grid_shape, image_shape, stack_shape, mask, density, *, coords=None, dcf=None):
        self.grid_shape = tuple(int(s) for s in grid_shape)
        self.image_shape = tuple(int(s) for s in image_shape)
        self.stack_shape = tuple(int(s) for s in stack_shape)
        self.mask = torch.as_tensor(mask).float()
        self.density = torch.as_tensor(density).float()
        self._coords = None if coords is None else torch.as_tensor(coords)
        self._dcf = None if dcf is None else torch.as_tensor(dcf).float()

    def __repr__(self):
        return (f"MaskedFFTPlan(grid_shape={self.grid_shape}, image_shape={self.image_shape}, "
                f"stack_shape={self.stack_shape}, mask={tuple(self.mask.shape)}, density={tuple(self.density.shape)})")


class MaskedFFT(SolveMixin):
    """Masked FFT / IFFT operator (reference _masked_fft.py:122-561)."""

    def __init__(self, grid_shape=None, image_shape=None, mask=None, density=None, smaps=None, stack_shape=None,
                 device=None, *, toeplitz=None, plan=None, coords=None, dcf=None):
        _plan_coords = None if coords is None else torch.as_tensor(coords)
        _plan_dcf = None if dcf is None else torch.as_tensor(dcf).float()
        if plan is not None and isinstance(plan, MaskedFFTPlan):
            grid_shape, image_shape, mask, density = plan.grid_shape, plan.image_shape, plan.mask, plan.density
            if stack_shape is None:
                stack_shape = plan.stack_shape
            if _plan_coords is None:
                _plan_coords = plan._coords
            if _plan_dcf is None:
                _plan_dcf = plan._dcf
        if grid_shape is None or image_shape is None or mask is None:
            raise ValueError(
                "Either 'plan' (MaskedFFTPlan) or explicit 'grid_shape', 'image_shape', and 'mask' are required."
            )
        comp_dev = device if (device is not None and torch.device(device).type == "cuda") else None
        self._comp_device = _lib.hip_device(comp_dev)
        self.grid_shape = tuple(int(s) for s in grid_shape)
        self.image_shape = tuple(int(s) for s in image_shape)
        self.ndim = len(self.grid_shape)
        self.grid_size = int(np.prod(self.grid_shape))
        self.fft_axes = tuple(range(-self.ndim, 0))
        mask_t = torch.as_tensor(np.asarray(mask) if isinstance(mask, np.ndarray) else mask).float()
        if stack_shape is None:
            stack_shape = tuple(int(s) for s in mask_t.shape[: -self.ndim]) if mask_t.ndim > self.ndim else ()
        self.stack_shape = tuple(int(s) for s in stack_shape)
        self.mask = mask_t
        self.density = None if density is None else torch.as_tensor(
            np.asarray(density) if isinstance(density, np.ndarray) else density).float()
        self._coords = _plan_coords
        self._dcf = _plan_dcf
        self.natural_shape = self.grid_shape
        self.n_samples = self.grid_size
        if smaps is not None:
            self.smaps = torch.as_tensor(smaps)
            self._conj_smaps = self.smaps.conj()
        else:
            self.smaps = None
            self._conj_smaps = None
        self._pad_slices = tuple(
            slice((gs - is_) // 2, (gs - is_) // 2 + is_) for gs, is_ in zip(self.grid_shape, self.image_shape)
        )
        self.device = torch.device(device) if device is not None else None
        if toeplitz is None:
            target = self.device if self.device is not None else torch.device("cpu")
            toeplitz = target.type == "cpu"
        if toeplitz and self.density is None:
            toeplitz = False
        self.toeplitz = bool(toeplitz)
        self._toep_op = None
        self._plan = plan
        self._mask_dev = None

    @property
    def n_coils(self):
        return None if self.smaps is None else int(self.smaps.shape[0])

    def _mask_s(self, S: int) -> torch.Tensor:
        if self._mask_dev is None:
            self._mask_dev = self.mask.to(self._comp_device).reshape(S, *self.grid_shape)
        return self._mask_dev

    def _signed_mask(self, S: int, half: bool, scale: float = 1.0) -> torch.Tensor:
        """mask * (-1)^(k0 + k1 + ...) (* prod (-1)^(N/2) when ``half``,
        * ``scale``): the k-space side of the centred transforms folded into
        the mask."""
        key = ("signed", half, float(scale))
        cache = self.__dict__.setdefault("_signed_cache", {})
        if key not in cache:
            m = self._mask_s(S).to(torch.float32)
            sg = grid_signs(self.grid_shape, m.device)
            if half:
                sg = sg * half_sign(self.grid_shape)
            w = m * sg
            if scale != 1.0:
                w = w * scale
            cache[key] = w.contiguous()
        return cache[key]

    def _stack_mask(self, s_flat_idx: int):
        if not self.stack_shape:
            return self.mask, self.density
        S = int(np.prod(self.stack_shape))
        m = self.mask.reshape(S, *self.grid_shape)[s_flat_idx]
        d = self.density.reshape(S, *self.grid_shape)[s_flat_idx] if self.density is not None else None
        return m, d

    def _split(self, prefix, what):
        s_shape = self.stack_shape
        if s_shape:
            if len(prefix) < len(s_shape) or tuple(prefix[-len(s_shape):]) != s_shape:
                raise ValueError(f"{what} prefix {prefix} must end with stack_shape {s_shape}")
            return prefix[: -len(s_shape)]
        return prefix

    @numpy_io
    def adjoint(self, kspace_grid: torch.Tensor) -> torch.Tensor:
        """(*B, *S, C, *grid_shape) -> (*B, *S, [C], *image_shape)."""
        x = torch.as_tensor(kspace_grid)
        prefix = tuple(int(s) for s in x.shape[: -(1 + self.ndim)])
        B_shape = self._split(prefix, "kspace_grid")
        src_device, dtype = x.device, x.dtype
        S = int(np.prod(self.stack_shape)) if self.stack_shape else 1
        B = int(np.prod(B_shape)) if B_shape else 1
        C = int(x.shape[-self.ndim - 1])
        d = self._comp_device
        ctype = dtype if dtype.is_complex else torch.complex64
        xd = x.to(d).to(ctype).reshape(B, S, C, *self.grid_shape)
        if ctype == torch.complex64 and even_grid(self.grid_shape):
            # signed mask, plain unnormalised IFFT, one pass for sign, scale,
            # crop and the conj(S) sum
            masked = xd * self._signed_mask(S, False).reshape(1, S, 1, *self.grid_shape)
            full = torch.fft.ifftn(masked.reshape(B * S, C, *self.grid_shape), dim=self.fft_axes, norm="forward")
            img = combine(full, self.smaps, B * S, C, self.image_shape, self.grid_shape,
                          half_sign(self.grid_shape) / math.sqrt(float(np.prod(self.grid_shape))))
            single = tuple(self.image_shape) if self.smaps is not None else (C, *self.image_shape)
            return img.reshape(*B_shape, *self.stack_shape, *single).to(src_device)
        m = self._mask_s(S).to(torch.float64 if ctype == torch.complex128 else torch.float32)
        masked = xd * m.reshape(1, S, 1, *self.grid_shape)
        img = resize(ifft(masked, axes=self.fft_axes), (B, S, C, *self.image_shape))
        if self.smaps is not None:
            img = (img * self._conj_smaps.to(d, ctype)).sum(2)
            single = tuple(self.image_shape)
        else:
            single = (C, *self.image_shape)
        return img.reshape(*B_shape, *self.stack_shape, *single).to(src_device)

    @numpy_io
    def forward(self, image: torch.Tensor) -> torch.Tensor:
        """(*B, *S, [C], *image_shape) -> (*B, *S, C, *grid_shape)."""
        x = torch.as_tensor(image)
        single_ndim = len(self.image_shape) + (0 if self.smaps is not None else 1)
        prefix = tuple(int(s) for s in x.shape[: x.ndim - single_ndim])
        B_shape = self._split(prefix, "image")
        src_device, dtype = x.device, x.dtype
        S = int(np.prod(self.stack_shape)) if self.stack_shape else 1
        B = int(np.prod(B_shape)) if B_shape else 1
        d = self._comp_device
        ctype = dtype if dtype.is_complex else torch.complex64
        xd = x.to(d).to(ctype)
        C = int(self.smaps.shape[0]) if self.smaps is not None else int(x.shape[-len(self.image_shape) - 1])
        if ctype == torch.complex64 and even_grid(self.grid_shape):
            # one pass for S_c x + zero pad + sign, plain unnormalised FFT,
            # signed mask carrying the 1/sqrt(N)
            padded = expand(xd.reshape(B * S, *(() if self.smaps is not None else (C,)), *self.image_shape),
                            self.smaps, B * S, C, self.image_shape, self.grid_shape)
            k = torch.fft.fftn(padded, dim=self.fft_axes, norm="backward").reshape(B, S, C, *self.grid_shape)
            scale = 1.0 / math.sqrt(float(np.prod(self.grid_shape)))
            k = k * self._signed_mask(S, True, scale).reshape(1, S, 1, *self.grid_shape)
            return k.reshape(*B_shape, *self.stack_shape, C, *self.grid_shape).to(src_device)
        if self.smaps is not None:
            sm = self.smaps.to(d, ctype)
            coil = xd.reshape(B, S, 1, *self.image_shape) * sm
        else:
            coil = xd.reshape(B, S, C, *self.image_shape)
        padded = torch.zeros(B, S, C, *self.grid_shape, dtype=ctype, device=d)
        padded[(slice(None), slice(None), slice(None), *self._pad_slices)] = coil
        k = fft(padded, axes=self.fft_axes)
        m = self._mask_s(S).to(torch.float64 if ctype == torch.complex128 else torch.float32)
        k = k * m.reshape(1, S, 1, *self.grid_shape)
        return k.reshape(*B_shape, *self.stack_shape, C, *self.grid_shape).to(src_device)

    def __call__(self, x, adjoint=False):
        return self.adjoint(x) if adjoint else self.forward(x)

    @numpy_io
    def normal(self, image: torch.Tensor) -> torch.Tensor:
        if self.toeplitz:
            if self._toep_op is None:
                from .._toep._grog_toep import GrogToeplitzOp

                self._toep_op = GrogToeplitzOp(self, device=self._comp_device)
            return self._toep_op(image)
        return self.adjoint(self.forward(image))

    def _mask_ifft_crop_batch(self, batch_kspace_grid: torch.Tensor, s_flat_idx: int = 0) -> torch.Tensor:
        x = torch.as_tensor(batch_kspace_grid)
        src = x.device
        R = int(x.shape[0])
        m, _ = self._stack_mask(s_flat_idx)
        d = self._comp_device
        k = x.to(d, torch.complex64) * m.to(d).unsqueeze(0)
        return resize(ifft(k, axes=self.fft_axes), (R, *self.image_shape)).to(src)

    def _fft_pad_mask_batch(self, batch_imgs: torch.Tensor, s_flat_idx: int = 0) -> torch.Tensor:
        x = torch.as_tensor(batch_imgs)
        src = x.device
        R = int(x.shape[0])
        m, _ = self._stack_mask(s_flat_idx)
        d = self._comp_device
        padded = torch.zeros(R, *self.grid_shape, dtype=torch.complex64, device=d)
        padded[(slice(None), *self._pad_slices)] = x.to(d, torch.complex64)
        return (fft(padded, axes=self.fft_axes) * m.to(d).unsqueeze(0)).to(src)
