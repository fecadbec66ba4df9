"""Sparse FFT operator for GROG-gridded data (drop-in for the reference's
src/pygrog/operator/_sparse_fft.py).

Naming follows the reference (S1): ``adjoint`` = sparse k-space -> image
(A^H), ``forward`` = image -> sparse k-space (A), ``normal`` = A^H A.

Where the reference loops over coils and stack elements in Python around a
C++ scatter / gather (:505-848), every call here is one device pass:
  * adjoint: one fused scatter kernel for all (batch, stack, coil) rows
    (``pg_grid_scatter``: the sort permutation is folded into the read, the
    per-cell sums are deterministic chunked folds, the dense grid is written
    once), then one batched centred IFFT and crop (torch.fft);
  * forward: one batched zero-pad + centred FFT, then one fused gather +
    unsort kernel (``pg_gather_perm_c64``).
Results come back on the input's device, as in the reference
(``accum.to(src_device)``).
"""

from __future__ import annotations

__all__ = ["SparseFFT", "gather", "scatter_add"]

import functools
import math

import numpy as np
import torch

from .. import _lib
from .._base._fftc import fft, ifft
from .._solve._mixin import SolveMixin
from .._utils import resize
from ._centred import combine, even_grid, expand, half_sign


def _get_torch_ext():
    """The ``_pygrog_torch`` op facade (reference _sparse_fft.py:44-117)."""
    return _lib.get_torch_ext()


def _scatter_add(grid, data, indices, weights, bin_starts=None, bin_size=0):
    """grid[indices[i]] += weights[i] * data[i], in place (reference :123-130)."""
    ext = _get_torch_ext()
    if bin_starts is not None:
        ext.scatter_add_binned(grid, data, indices, weights, bin_starts, bin_size)
    else:
        ext.scatter_add(grid, data, indices, weights)


def _gather(grid, indices, weights):
    """out[i] = weights[i] * grid[indices[i]] (reference :133-135)."""
    return _get_torch_ext().gather(grid, indices, weights)


def scatter_add(grid: torch.Tensor, data: torch.Tensor, indices: torch.Tensor, weights: torch.Tensor) -> None:
    """Scatter-add ``grid[indices[i]] += weights[i] * data[i]`` (in place)."""
    _scatter_add(grid, data, indices, weights)


def gather(grid: torch.Tensor, indices: torch.Tensor, weights: torch.Tensor) -> torch.Tensor:
    """Gather ``out[i] = weights[i] * grid[indices[i]]``."""
    return _gather(grid, indices, weights)


def numpy_io(fn):
    """numpy in -> numpy out, torch in -> torch out (the role of the
    reference's ``mrinufft.with_torch`` on these methods)."""

    @functools.wraps(fn)
    def wrapper(self, x, *args, **kwargs):
        is_np = isinstance(x, np.ndarray)
        out = fn(self, torch.as_tensor(x) if is_np else x, *args, **kwargs)
        if is_np and isinstance(out, torch.Tensor):
            return out.cpu().numpy()
        return out

    return wrapper


class _ScatterPlan:
    """Device-resident scatter structure of a SparseFFT (per stack element:
    CSR over cells, chunked reduction units, permutations)."""

    def __init__(self, indices, sqrt_w, sort_perm, inv_perm, grid_size: int, S: int, n: int, device):
        from ..calib._grog import HOT_MIN_UNITS, UNIT_CHUNK, compile_units, hot_lists, ordinary_rows

        self.device = device
        self.S, self.n, self.G = S, n, grid_size
        st = _lib.stream(device)
        idx = indices.reshape(S, n).to(device, torch.int64).contiguous()
        sqw = sqrt_w.reshape(S, n).to(device, torch.float32).contiguous()
        self.sort_perm = sort_perm.reshape(S, n).to(device, torch.int64).contiguous()
        self.inv_perm = inv_perm.reshape(S, n).to(device, torch.int64).contiguous()
        if n > 0:
            bad = (idx < 0) | (idx >= grid_size)
            if bool(bad.any()):
                # out-of-range samples are skipped by scatter and read as 0 by
                # gather (reference sparse_ops_cpu_impl.inl:98, 161): keep
                # them, sorted, with zero weight on a clamped index
                idx = idx.clamp(0, max(grid_size - 1, 0))
                sqw = torch.where(bad, torch.zeros_like(sqw), sqw)
        self.indices = idx
        self.sqrt_w = sqw
        self.cell_ptr = torch.empty(S, grid_size + 1, dtype=torch.int32, device=device)
        for s in range(S):
            _lib.call("pg_csr_from_sorted", idx[s].data_ptr(), n, grid_size, self.cell_ptr[s].data_ptr(), st)
        self.cell_unit, self.unit_cell, self.unit_rep = compile_units(self.cell_ptr, S, grid_size, device, UNIT_CHUNK)
        min_units = HOT_MIN_UNITS
        self.unit_hot, self.hot = hot_lists(self.cell_unit, S, grid_size, device, min_units)
        self.cell_vrow = ordinary_rows(self.cell_ptr, self.cell_unit, S, grid_size, min_units)
        self.rep_src = self.sort_perm.to(torch.int32).contiguous()
        self._signed = {}

    def signed_weights(self, grid_shape, half: bool, scale: float = 1.0) -> torch.Tensor:
        """sqrt weights times (-1)^(k0 + k1 + ...) of each sample's cell
        (times prod (-1)^(N/2) when ``half``, times ``scale``): the k-space
        side of the centred transforms, folded into the scatter / gather
        weights."""
        key = (half, float(scale))
        if key not in self._signed:
            from ._centred import cell_signs, half_sign

            sg = cell_signs(self.indices, grid_shape)
            if half:
                sg = sg * half_sign(grid_shape)
            w = self.sqrt_w * sg
            if scale != 1.0:
                w = w * scale
            self._signed[key] = w.contiguous()
        return self._signed[key]

    def frames(self, y_ptr: int, grid_ptr: int, B: int, C: int, y_stride_s: int | None = None,
               rep_s: torch.Tensor | None = None):
        """Frames for data (B, S, C, n) -> grid (B, S, C, G)."""
        fr = []
        S, n, G = self.S, self.n, self.G
        for b in range(B):
            for s in range(S):
                f = _lib.Frame()
                f.y = y_ptr + ((b * S + s) * C * n) * 8
                f.grid = grid_ptr + ((b * S + s) * C * G) * 8
                f.rep_src = self.rep_src.data_ptr() + s * n * 4
                f.rep_s = (self.sqrt_w if rep_s is None else rep_s).data_ptr() + s * n * 4
                f.unit_cell = self.unit_cell.data_ptr()
                f.unit_rep = self.unit_rep.data_ptr() + s * 4
                f.cell_unit = self.cell_unit.data_ptr() + s * G * 4
                hot = self.hot[s]
                f.unit_hot = self.unit_hot.data_ptr()
                f.hot_units = hot["hot_units"].data_ptr()
                f.hot_ptr = hot["hot_ptr"].data_ptr()
                f.n_hot = hot["n_hot"]
                f.n_hot_units = hot["n_hot_units"]
                f.cell_row = self.cell_ptr[s].data_ptr()
                f.cell_vrow = self.cell_vrow[s].data_ptr()
                f.n_src = n
                f.n_cell = G
                fr.append(f)
        return fr

    def scatter(self, data: torch.Tensor, rep_s: torch.Tensor | None = None) -> torch.Tensor:
        """data (B, S, C, n) complex64 on device -> grid (B, S, C, G);
        ``rep_s`` replaces the sqrt weights (signed weights)."""
        from ..calib._grog import GridProgram

        B, S, C, n = data.shape
        out = torch.empty(B, S, C, self.G, dtype=torch.complex64, device=self.device)
        if C == 0 or B == 0:
            return out
        for c0 in range(0, C, 64):
            c1 = min(C, c0 + 64)
            if c0 == 0 and c1 == C:
                src, dst = data, out
            else:
                src = data[:, :, c0:c1].contiguous()
                dst = torch.empty(B, S, c1 - c0, self.G, dtype=torch.complex64, device=self.device)
            tcc = _lib.load().pg_grid_tile_cells(c1 - c0)
            prog = GridProgram(self.frames(src.data_ptr(), dst.data_ptr(), B, c1 - c0, rep_s=rep_s), c1 - c0,
                               self.G, self.device, tcc)
            prog.launch("pg_grid_scatter")
            if dst is not out:
                out[:, :, c0:c1] = dst
        return out

    def gather(self, kgrid: torch.Tensor, weights: torch.Tensor | None = None) -> torch.Tensor:
        """kgrid (B, S, C, G) -> (B, S, C, n) natural order, sqrt_w (or
        ``weights``, signed) applied."""
        B, S, C, G = kgrid.shape
        out = torch.empty(B, S, C, self.n, dtype=torch.complex64, device=self.device)
        kg = kgrid.contiguous()
        w = (self.sqrt_w if weights is None else weights).contiguous()
        idx, inv = self.indices.contiguous(), self.inv_perm.contiguous()
        # every (batch, stack element, coil) row in one launch
        _lib.call("pg_gather_perm_stack_c64", _lib.ptr(kg), G, _lib.ptr(idx), _lib.ptr(w), _lib.ptr(inv), self.n, C,
                  S, B, _lib.ptr(out), _lib.stream(self.device))
        return out


class SparseFFT(SolveMixin):
    """Sparse FFT / IFFT operator with coil combination (reference
    ``SparseFFT``, _sparse_fft.py:184-1242).

    ``device`` keeps the reference's meaning for the Toeplitz auto policy
    (``toeplitz=None`` -> on unless ``device`` is a GPU); the arithmetic
    always runs on a HIP device (the given one, else the current one).
    """

    def __init__(self, grid_shape=None, image_shape=None, indices=None, weights=None, smaps=None, device=None, *,
                 plan=None, toeplitz=None):
        if plan is None and (grid_shape is None or image_shape is None):
            raise ValueError("Either 'plan' or both 'grid_shape'/'image_shape' required")
        comp_dev = device if (device is not None and torch.device(device).type == "cuda") else None
        self._comp_device = _lib.hip_device(comp_dev)
        dev_fields = None
        if plan is not None:
            self.grid_shape = tuple(int(s) for s in plan.grid_shape)
            self.image_shape = tuple(int(s) for s in plan.image_shape)
            self.grid_size = int(plan.grid_size)
            dev_fields = plan.__dict__.get("_dev")
            self.natural_shape = tuple(int(s) for s in getattr(plan, "natural_shape", (int(plan.n_samples),)))
            self.stack_shape = tuple(int(s) for s in (getattr(plan, "stack_shape", ()) or ()))
            self._plan = plan
        else:
            if grid_shape is None or image_shape is None:
                raise ValueError("Either 'plan' or both 'grid_shape'/'image_shape' required")
            self.grid_shape = tuple(int(s) for s in grid_shape)
            self.image_shape = tuple(int(s) for s in image_shape)
            self.grid_size = int(np.prod(self.grid_shape))
            idx = torch.as_tensor(np.asarray(indices) if not isinstance(indices, torch.Tensor) else indices)
            idx = idx.reshape(-1).to(torch.int64)
            w = torch.as_tensor(np.asarray(weights) if not isinstance(weights, torch.Tensor) else weights)
            w = w.reshape(-1).to(torch.float32)
            home = idx.device
            d = self._comp_device
            idx_d, w_d = idx.to(d), w.to(d)
            sqrt_w = torch.sqrt(w_d)
            # stable sort of the indices (reference uses torch.argsort)
            sort_perm = torch.argsort(idx_d, stable=True)
            inv_perm = torch.empty_like(sort_perm)
            inv_perm[sort_perm] = torch.arange(sort_perm.numel(), device=d)
            self._plan = None
            self._home_fields = {
                "indices": idx_d[sort_perm].to(home),
                "sqrt_weights": sqrt_w[sort_perm].to(home),
                "sort_perm": sort_perm.to(home),
                "inv_perm": inv_perm.to(home),
            }
            dev_fields = {"indices": idx_d[sort_perm], "sqrt_weights": sqrt_w[sort_perm], "sort_perm": sort_perm,
                          "inv_perm": inv_perm}
            self.natural_shape = (int(idx.numel()),)
            self.stack_shape = ()
        self._dev_fields = dev_fields
        self.n_samples = int(self._field("indices").shape[-1])
        self.ndim = len(self.grid_shape)
        self.fft_axes = tuple(range(-self.ndim, 0))

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
        self.toeplitz = bool(toeplitz)
        self._toep_op = None
        self._scatter_plan = None

    # -- reference attribute surface -------------------------------------------
    def _field(self, name):
        if self._dev_fields is not None and self._dev_fields.get(name) is not None:
            return self._dev_fields[name]
        return torch.as_tensor(getattr(self._plan, name))

    @property
    def indices(self):
        return self._home_fields["indices"] if self._plan is None else self._plan.indices

    @property
    def sqrt_weights(self):
        return self._home_fields["sqrt_weights"] if self._plan is None else self._plan.sqrt_weights

    @property
    def sort_perm(self):
        return self._home_fields["sort_perm"] if self._plan is None else self._plan.sort_perm

    @property
    def inv_perm(self):
        return self._home_fields["inv_perm"] if self._plan is None else self._plan.inv_perm

    @property
    def n_coils(self):
        return None if self.smaps is None else int(self.smaps.shape[0])

    def _sp(self) -> _ScatterPlan:
        if self._scatter_plan is None:
            S = int(np.prod(self.stack_shape)) if self.stack_shape else 1
            self._scatter_plan = _ScatterPlan(self._field("indices"), self._field("sqrt_weights"),
                                              self._field("sort_perm"), self._field("inv_perm"), self.grid_size, S,
                                              self.n_samples, self._comp_device)
        return self._scatter_plan

    def _split_prefix(self, prefix, what):
        s_shape = self.stack_shape
        if s_shape:
            if len(prefix) < len(s_shape) or tuple(prefix[-len(s_shape):]) != s_shape:
                raise ValueError(f"{what} prefix {prefix} must end with stack_shape {s_shape}")
            return prefix[: -len(s_shape)]
        return prefix

    # -- A^H: sparse k-space -> image -------------------------------------------
    @numpy_io
    def adjoint(self, sparse_kspace: torch.Tensor) -> torch.Tensor:
        """(*B, *S, C, n_samples | *natural_shape) -> (*B, *S, [C], *image_shape)."""
        x = torch.as_tensor(sparse_kspace)
        nat = self.natural_shape
        if len(nat) > 1 and tuple(int(s) for s in x.shape[-len(nat):]) == nat:
            x = x.reshape(*x.shape[: -len(nat)], self.n_samples)
        prefix = tuple(int(s) for s in x.shape[:-2])
        B_shape = self._split_prefix(prefix, "sparse_kspace")
        src_device, dtype = x.device, x.dtype
        S = int(np.prod(self.stack_shape)) if self.stack_shape else 1
        B = int(np.prod(B_shape)) if B_shape else 1
        C = int(x.shape[-2])
        sp = self._sp()
        xd = x.to(sp.device).to(torch.complex64).reshape(B, S, C, self.n_samples).contiguous()
        if even_grid(self.grid_shape):
            # scatter with (-1)^k-signed weights, plain IFFT, then one pass for
            # the (-1)^(n - M) sign, the crop and the conj(S) coil sum
            # the IFFT runs unnormalised: its 1/sqrt(N) rides on that pass too
            g = sp.scatter(xd, rep_s=sp.signed_weights(self.grid_shape, False))
            full = torch.fft.ifftn(g.reshape(B * S, C, *self.grid_shape), dim=self.fft_axes, norm="forward")
            img = combine(full, self.smaps, B * S, C, self.image_shape, self.grid_shape,
                          half_sign(self.grid_shape) / math.sqrt(self.grid_size))
            single = tuple(self.image_shape) if self.smaps is not None else (C, *self.image_shape)
        else:
            g = sp.scatter(xd)
            img = ifft(g.reshape(B, S, C, *self.grid_shape), axes=self.fft_axes)
            img = resize(img, (B, S, C, *self.image_shape))
            if self.smaps is not None:
                cs = self._conj_smaps.to(sp.device, torch.complex64)
                img = (img * cs).sum(2)
                single = tuple(self.image_shape)
            else:
                single = (C, *self.image_shape)
        out = img.reshape(*B_shape, *self.stack_shape, *single)
        if out.dtype != dtype and dtype.is_complex:
            out = out.to(dtype)
        return out.to(src_device)

    # -- A: image -> sparse k-space -----------------------------------------------
    @numpy_io
    def forward(self, image: torch.Tensor) -> torch.Tensor:
        """(*B, *S, [C], *image_shape) -> (*B, *S, C, *natural_shape)."""
        out = self._adjoint_flat(image)
        nat = self.natural_shape
        if len(nat) > 1:
            out = out.reshape(*out.shape[:-1], *nat)
        return out

    def _adjoint_flat(self, image: torch.Tensor) -> torch.Tensor:
        x = torch.as_tensor(image)
        single_ndim = len(self.image_shape) + (0 if self.smaps is not None else 1)
        prefix = tuple(int(s) for s in x.shape[: x.ndim - single_ndim])
        B_shape = self._split_prefix(prefix, "image")
        src_device, dtype = x.device, x.dtype
        S = int(np.prod(self.stack_shape)) if self.stack_shape else 1
        B = int(np.prod(B_shape)) if B_shape else 1
        sp = self._sp()
        xd = x.to(sp.device)
        if not xd.is_complex():
            xd = xd.to(torch.complex64)
        xd = xd.to(torch.complex64)
        if self.smaps is not None:
            C = int(self.smaps.shape[0])
        else:
            C = int(x.shape[-len(self.image_shape) - 1])
        if even_grid(self.grid_shape):
            # one pass for S_c x, the zero pad and the (-1)^n sign; plain,
            # unnormalised FFT; the (-1)^(k - M) sign and the 1/sqrt(N) ride on
            # the gather weights
            padded = expand(xd.reshape(B * S, *(() if self.smaps is not None else (C,)), *self.image_shape),
                            self.smaps, B * S, C, self.image_shape, self.grid_shape)
            k = torch.fft.fftn(padded, dim=self.fft_axes, norm="backward").reshape(B, S, C, self.grid_size)
            out = sp.gather(k, sp.signed_weights(self.grid_shape, True, 1.0 / math.sqrt(self.grid_size)))
        else:
            if self.smaps is not None:
                sm = self.smaps.to(sp.device, torch.complex64)
                coil = xd.reshape(B, S, 1, *self.image_shape) * sm
            else:
                coil = xd.reshape(B, S, C, *self.image_shape)
            padded = torch.zeros(B, S, C, *self.grid_shape, dtype=torch.complex64, device=sp.device)
            padded[(slice(None), slice(None), slice(None), *self._pad_slices)] = coil
            k = fft(padded, axes=self.fft_axes).reshape(B, S, C, self.grid_size)
            out = sp.gather(k)
        out = out.reshape(*B_shape, *self.stack_shape, C, self.n_samples)
        if dtype.is_complex and out.dtype != dtype:
            out = out.to(dtype)
        return out.to(src_device)

    def __call__(self, x, adjoint=False):
        return self.adjoint(x) if adjoint else self.forward(x)

    # -- A^H A ------------------------------------------------------------------
    @numpy_io
    def normal(self, image: torch.Tensor) -> torch.Tensor:
        """``A^H A image``: Toeplitz PSF embedding when ``self.toeplitz``,
        else the composition (reference :858-887)."""
        if self.toeplitz:
            if self._toep_op is None:
                from .._toep._grog_toep import GrogToeplitzOp

                self._toep_op = GrogToeplitzOp(self, device=self._comp_device)
            return self._toep_op(image)
        return self.adjoint(self.forward(image))

    # -- batch helpers (reference :1050-1149) -----------------------------------
    def _scatter_ifft_crop_batch(self, batch_kspace: torch.Tensor, s_flat_idx: int = 0) -> torch.Tensor:
        """(R, n_samples) unsorted rows of one stack element -> (R, *image_shape)."""
        x = torch.as_tensor(batch_kspace)
        src_device = x.device
        R = int(x.shape[0])
        sp = self._sp()
        S = sp.S
        xd = x.to(sp.device, torch.complex64)
        if S > 1:
            full = torch.zeros(1, S, R, self.n_samples, dtype=torch.complex64, device=sp.device)
            full[0, s_flat_idx] = xd
            g = sp.scatter(full)[0, s_flat_idx]
        else:
            g = sp.scatter(xd.reshape(1, 1, R, self.n_samples).contiguous())[0, 0]
        img = ifft(g.reshape(R, *self.grid_shape), axes=self.fft_axes)
        return resize(img, (R, *self.image_shape)).to(src_device)

    def _fft_pad_gather_batch(self, batch_imgs: torch.Tensor, s_flat_idx: int = 0) -> torch.Tensor:
        """(R, *image_shape) -> (R, n_samples) in natural order."""
        x = torch.as_tensor(batch_imgs)
        src_device = x.device
        R = int(x.shape[0])
        sp = self._sp()
        padded = torch.zeros(R, *self.grid_shape, dtype=torch.complex64, device=sp.device)
        padded[(slice(None), *self._pad_slices)] = x.to(sp.device, torch.complex64)
        k = fft(padded, axes=self.fft_axes).reshape(R, self.grid_size).contiguous()
        out = torch.empty(R, self.n_samples, dtype=torch.complex64, device=sp.device)
        _lib.call("pg_gather_perm_c64", _lib.ptr(k), self.grid_size, sp.indices[s_flat_idx].data_ptr(),
                  sp.sqrt_w[s_flat_idx].data_ptr(), sp.inv_perm[s_flat_idx].data_ptr(), self.n_samples, R,
                  _lib.ptr(out), _lib.stream(sp.device))
        return out.to(src_device)
