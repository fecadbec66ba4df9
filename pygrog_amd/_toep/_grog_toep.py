"""GROG-Toeplitz normal operator (drop-in for the reference's
src/pygrog/_toep/_grog_toep.py).

``A^H A x = sum_c conj(S_c) crop(IFFT(PSF * FFT(pad(S_c x))))``, with the PSF
on the (oversampled) GROG grid: no 2x padding, because GROG samples sit on
the grid.  PSF = sum of squared sqrt-weights per cell: the density of a
MaskedFFT or of the compiled plan behind a plan-backed SparseFFT, otherwise
built once by the order-fixed ``psf_scatter_scalar`` HIP kernel.  All coils (and stack elements) go through one batched
FFT / multiply / IFFT on the GPU.
"""

from __future__ import annotations

__all__ = ["GrogToeplitzOp"]

import numpy as np
import torch

from .. import _lib
from .._base._fftc import fft, ifft
from .._utils import resize
from ..operator._centred import combine, even_grid, expand


class GrogToeplitzOp:
    """Self-adjoint ``A^H A`` of a SparseFFT / MaskedFFT (reference :30-209)."""

    def __init__(self, sparse_fft, device=None):
        self.sparse_fft = sparse_fft
        self.device = _lib.hip_device(device if (device is not None and torch.device(device).type == "cuda") else None)
        self.grid_shape = tuple(sparse_fft.grid_shape)
        self.image_shape = tuple(sparse_fft.image_shape)
        self.fft_axes = sparse_fft.fft_axes
        self._pad_slices = sparse_fft._pad_slices
        self.stack_shape = tuple(getattr(sparse_fft, "stack_shape", ()) or ())
        S = int(np.prod(self.stack_shape)) if self.stack_shape else 1
        comp = getattr(getattr(sparse_fft, "_plan", None), "_compiled", None)
        if getattr(sparse_fft, "density", None) is not None:
            dens = torch.as_tensor(sparse_fft.density).to(self.device).float()
            psf = dens.reshape(S, *self.grid_shape).clone()
        elif comp is not None and getattr(comp, "density", None) is not None and comp.density.numel() == S * int(
                np.prod(self.grid_shape)):
            # a compiled GROG plan already holds sum(sqrt_w^2) per cell, folded
            # in the plan's sorted order from zero: the scalar PSF, bit for bit
            psf = comp.density.to(self.device).reshape(S, *self.grid_shape).clone()
        else:
            ext = _lib.get_torch_ext()
            G = int(np.prod(self.grid_shape))
            sqw = sparse_fft._field("sqrt_weights").to(self.device).reshape(S, -1)
            idx = sparse_fft._field("indices").to(self.device).reshape(S, -1)
            psf = torch.zeros(S, G, dtype=sqw.dtype, device=self.device)
            for s in range(S):
                w_sq = (sqw[s] * sqw[s]).contiguous()
                ext.psf_scatter_scalar(psf[s], idx[s].contiguous(), w_sq)
            psf = psf.reshape(S, *self.grid_shape)
        self.psf = psf if self.stack_shape else psf[0]
        self._psf_c = psf.to(torch.complex64 if psf.dtype == torch.float32 else torch.complex128)
        # the fast path runs both transforms unnormalised and applies their
        # 1/N here, in the one multiply between them
        self._psf_n = (psf.to(torch.float32) / float(np.prod(self.grid_shape))).to(torch.complex64)

    def __call__(self, image: torch.Tensor) -> torch.Tensor:
        sf = self.sparse_fft
        x = torch.as_tensor(image)
        single_ndim = len(self.image_shape) + (0 if sf.smaps is not None else 1)
        prefix = tuple(int(s) for s in x.shape[: x.ndim - single_ndim])
        s_shape = self.stack_shape
        if s_shape:
            if len(prefix) < len(s_shape) or tuple(prefix[-len(s_shape):]) != s_shape:
                raise ValueError(f"image prefix {prefix} must end with stack_shape {s_shape}")
            B_shape = prefix[: -len(s_shape)]
        else:
            B_shape = prefix
        S = int(np.prod(s_shape)) if s_shape else 1
        B = int(np.prod(B_shape)) if B_shape else 1
        src, dtype = x.device, x.dtype
        ctype = dtype if dtype.is_complex else torch.complex64
        xd = x.to(self.device).to(ctype)
        C = int(sf.smaps.shape[0]) if sf.smaps is not None else int(x.shape[-len(self.image_shape) - 1])
        if ctype == torch.complex64 and even_grid(self.grid_shape):
            # (-1)^n sign + pad in the expand pass; the k-space signs of the two
            # centred transforms cancel around the real PSF, which also carries
            # both transforms' scale; crop + (-1)^n + conj(S) coil sum in the
            # combine pass
            padded = expand(xd.reshape(B * S, *(() if sf.smaps is not None else (C,)), *self.image_shape),
                            sf.smaps, B * S, C, self.image_shape, self.grid_shape)
            k = torch.fft.fftn(padded, dim=self.fft_axes, norm="backward").reshape(B, S, C, *self.grid_shape)
            k *= self._psf_n.reshape(1, S, 1, *self.grid_shape)
            full = torch.fft.ifftn(k.reshape(B * S, C, *self.grid_shape), dim=self.fft_axes, norm="forward")
            out = combine(full, sf.smaps, B * S, C, self.image_shape, self.grid_shape, 1.0)
            single = tuple(self.image_shape) if sf.smaps is not None else (C, *self.image_shape)
            return out.reshape(*B_shape, *s_shape, *single).to(src)
        psf = self._psf_c.to(ctype).reshape(1, S, 1, *self.grid_shape)
        if sf.smaps is not None:
            sm = sf.smaps.to(self.device, ctype)
            coil = xd.reshape(B, S, 1, *self.image_shape) * sm
        else:
            coil = xd.reshape(B, S, C, *self.image_shape)
        if self.grid_shape == self.image_shape:
            full = ifft(fft(coil, axes=self.fft_axes) * psf, axes=self.fft_axes)
        else:
            padded = torch.zeros(B, S, C, *self.grid_shape, dtype=ctype, device=self.device)
            padded[(slice(None), slice(None), slice(None), *self._pad_slices)] = coil
            full = ifft(fft(padded, axes=self.fft_axes) * psf, axes=self.fft_axes)
            full = resize(full, (B, S, C, *self.image_shape))
        if sf.smaps is not None:
            out = (full * sf._conj_smaps.to(self.device, ctype)).sum(2)
            single = tuple(self.image_shape)
        else:
            out = full
            single = (C, *self.image_shape)
        return out.reshape(*B_shape, *s_shape, *single).to(src)
