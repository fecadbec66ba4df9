"""Loader of ``libpygrog_hip.so`` and the ``_pygrog_torch``-compatible op facade.

The reference binds seven ops through a pybind11 module ``_pygrog_torch``
(csrc/torch/module.cpp:54-87), returned by ``_get_torch_ext()``
(src/pygrog/calib/_grog.py:191-264, src/pygrog/operator/_sparse_fft.py:44-117).
Here the same names and argument meanings are served by :class:`TorchExt`,
whose methods validate shapes and dtypes with the reference's messages and then
call the C ABI of ``include/pygrog_hip.h`` on HIP device tensors, enqueued on
torch's current stream; ``get_torch_ext()`` returns them registered as torch
custom operators (``pygrog_amd._torch_ops``).

There is no CPU implementation behind these ops: if the shared library is
missing, or a tensor is not on a HIP device, the call raises.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import threading

import torch

_LIB_PATH = pathlib.Path(__file__).resolve().parent / "_lib" / "libpygrog_hip.so"

c_void_p = ctypes.c_void_p
c_int64 = ctypes.c_int64
c_int32 = ctypes.c_int32
c_float = ctypes.c_float
c_size_t_p = ctypes.POINTER(ctypes.c_size_t)


class PlanGeom(ctypes.Structure):
    """Mirror of ``pg_plan_geom`` (include/pygrog_hip.h)."""

    _fields_ = [
        ("ndim", c_int32),
        ("kw", c_int32),
        ("grid_shape", c_int64 * 3),
        ("strides", c_int64 * 3),
        ("half_amp", c_float * 3),
        ("origins", c_float * 3),
        ("steps", c_float * 3),
        ("offsets", (c_int32 * 3) * 128),
    ]


class Frame(ctypes.Structure):
    """Mirror of ``pg_frame`` (include/pygrog_hip.h)."""

    _fields_ = [
        ("y", c_void_p),
        ("grid", c_void_p),
        ("table", c_void_p),
        ("rep_src", c_void_p),
        ("rep_ent", c_void_p),
        ("rep_s", c_void_p),
        ("unit_cell", c_void_p),
        ("unit_rep", c_void_p),
        ("cell_unit", c_void_p),
        ("vbuf", c_void_p),
        ("unit_hot", c_void_p),
        ("hot_units", c_void_p),
        ("hot_ptr", c_void_p),
        ("p_unit", c_void_p),
        ("cell_row", c_void_p),
        ("cell_vrow", c_void_p),
        ("n_src", c_int64),
        ("n_cell", c_int64),
        ("n_hot", c_int64),
        ("n_hot_units", c_int64),
    ]


FRAME_BYTES = ctypes.sizeof(Frame)

# name -> argument types (restype int unless listed in _RESTYPES)
_SIGS = {
    "pg_last_error": [],
    "pg_abi_version": [],
    "pg_grog_interpolate_c64": [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_void_p],
    "pg_grog_interpolate_c128": [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_void_p],
    "pg_scatter_add_c64": [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_scatter_add_c128": [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_scatter_add_rows_c64": [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_void_p],
    "pg_scatter_add_rows_c128": [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_void_p],
    "pg_scatter_add_binned_c64": [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int64, c_void_p],
    "pg_scatter_add_binned_c128": [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_int64, c_void_p],
    "pg_gather_c64": [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p, c_void_p],
    "pg_gather_c128": [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p, c_void_p],
    "pg_psf_scatter_scalar_f32": [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_psf_scatter_scalar_f64": [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_psf_scatter_outer_c64": [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_psf_scatter_outer_c128": [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_psf_scatter_outer_basis_c64": [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p],
    "pg_psf_scatter_outer_basis_c128": [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p],
    "pg_plan_absmax": [c_void_p, c_int64, c_int64, c_int32, c_void_p, c_void_p],
    "pg_plan_targets": [c_void_p, c_void_p, c_int64, c_int64, PlanGeom, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "pg_sort_pairs_u64": [c_void_p, c_size_t_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_void_p],
    "pg_sort_pairs_u32": [c_void_p, c_size_t_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_void_p],
    "pg_exclusive_scan_u32": [c_void_p, c_size_t_p, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_plan_post": [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "pg_plan_weights": [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_plan_units": [c_void_p, c_void_p, c_int64, c_int64, c_int32, c_void_p, c_void_p, c_void_p],
    "pg_grappa_gram_f64": [c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p],
    "pg_plan_hot_counts": [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p],
    "pg_plan_hot_rows": [c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                         c_void_p, c_void_p],
    "pg_plan_quantize": [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int32, c_float, c_float, c_float, c_int64, c_int64, c_void_p, c_void_p, c_int64, c_void_p],
    "pg_table_pack_c64": [c_void_p, c_int64, c_int64, c_void_p, c_void_p],
    "pg_table_packed_elems": [c_int64, c_int64],
    "pg_grid_fp32": [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_void_p],
    "pg_grid_reduce_v": [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_void_p],
    "pg_grid_reduce_v_bfp": [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64,
                             c_void_p],
    "pg_grid_scatter": [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_int64, c_void_p],
    "pg_grid_tile_cells": [c_int64],
    "pg_grid_tile_options": [c_int64, c_void_p, ctypes.c_int32],
    "pg_gather_perm_c64": [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_void_p, c_void_p],
    "pg_gather_perm_stack_c64": [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64,
                                 c_void_p, c_void_p],
    "pg_unit_count": [c_void_p, c_int64, c_int64, c_int32, c_void_p, c_void_p],
    "pg_transpose_cn_c64": [c_void_p, c_int64, c_int64, c_void_p, c_void_p],
    "pg_transpose_frames_c64": [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p],
    "pg_s24_row_bytes": [c_int64],
    "pg_transpose_cn_s24": [c_void_p, c_int64, c_int64, c_void_p, c_void_p],
    "pg_transpose_frames_s24": [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p],
    "pg_bfp_row_bytes": [c_int64],
    "pg_transpose_cn_bfp": [c_void_p, c_int64, c_int64, c_void_p, c_void_p],
    "pg_transpose_frames_bfp": [c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p],
    "pg_vdot_c64_f64": [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p],
    "pg_vdot_c64_f64_noalloc": [c_void_p, c_void_p, c_int64, c_void_p, c_void_p],
    "pg_plan_compile_layout": [c_int64, c_int64, c_int32, c_int32, c_int64, c_void_p, c_void_p, c_void_p],
    "pg_plan_compile": [c_void_p, c_int64, c_int64, PlanGeom, c_int64, c_int32, c_int64, c_void_p, c_int64, c_void_p,
                        c_int64, c_void_p, c_void_p, c_void_p],
    "pg_plan_compile_async": [c_void_p, c_int64, c_int64, PlanGeom, c_int64, c_int32, c_int64, c_void_p, c_int64,
                              c_void_p, c_int64, c_void_p, c_void_p, c_void_p],
    "pg_vdot_workspace_bytes": [],
    "pg_sense_expand_c64": [c_void_p, c_void_p, c_int64, c_int64, ctypes.c_int32, c_void_p, c_void_p,
                            ctypes.c_int32, c_void_p, c_void_p],
    "pg_sense_combine_c64": [c_void_p, c_void_p, c_int64, c_int64, ctypes.c_int32, c_void_p, c_void_p,
                             ctypes.c_int32, ctypes.c_float, c_void_p, c_void_p],
    "pg_table_pack_bf16": [c_void_p, c_int64, c_int64, c_int32, c_void_p, c_void_p],
    "pg_table_bf16_elems": [c_int64, c_int64, c_int32],
    "pg_cg_step_xr": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p],
    "pg_cg_step_p": [c_void_p, c_void_p, c_int64, c_void_p, c_void_p],
    "pg_group_by_entry_layout": [c_int64, c_int64, c_int64, c_void_p, c_void_p],
    "pg_group_by_entry": [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p, c_void_p,
                          c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p],
    "pg_table_compose_f64": [c_void_p, c_int32, c_int64, c_int64, c_int32, c_void_p, c_void_p, c_void_p],
    "pg_group_by_entry_async": [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p],
    "pg_apply_grouped_bf16": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int32, c_void_p, c_void_p],
    "pg_apply_grouped_bf16_s24": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int32, c_void_p, c_void_p],
    "pg_apply_grouped_bf16_bfp": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int32, c_void_p, c_void_p],
    "pg_sparse_apply_c64": [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64, c_void_p],
    "pg_sparse_apply_stack_c64": [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_int64, c_int64, c_int64,
                                  c_int64, c_void_p],
    "pg_csr_from_sorted": [c_void_p, c_int64, c_int64, c_void_p, c_void_p],
    "pg_csr_from_sorted_u32": [c_void_p, c_int64, c_int64, c_void_p, c_void_p],
}
_RESTYPES = {
    "pg_last_error": ctypes.c_char_p,
    "pg_table_bf16_elems": c_int64,
    "pg_table_packed_elems": c_int64,
    "pg_s24_row_bytes": c_int64,
    "pg_bfp_row_bytes": c_int64,
    "pg_vdot_workspace_bytes": c_int64,
}

_lib = None
_lock = threading.Lock()


def lib_path() -> pathlib.Path:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not _LIB_PATH.exists():
            raise RuntimeError(
                f"pygrog_amd: HIP library not found at {_LIB_PATH}. Build it with "
                "`python -m pygrog_amd.csrc.build` (needs hipcc, gfx950)."
            )
        lib = ctypes.CDLL(str(_LIB_PATH))
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = lib
        return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def call(name: str, *args) -> None:
    """Call a C-ABI entry point and raise RuntimeError on a nonzero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.pg_last_error().decode(errors="replace")
        raise RuntimeError(f"{name}: {msg}")


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    return t.data_ptr()


def tile_options(C: int) -> list:
    """Tile sizes (cells) compiled for ``C`` coils, largest first."""
    buf = (ctypes.c_int32 * 8)()
    n = load().pg_grid_tile_options(int(C), buf, 8)
    return [int(buf[i]) for i in range(min(n, 8))]


def dims(shape) -> "ctypes.Array":
    """Host int64 array of a shape, for the entries that take one."""
    return (ctypes.c_int64 * 3)(*[int(v) for v in shape], *([1] * (3 - len(shape))))


def stream(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_hip(*tensors: torch.Tensor) -> torch.device:
    """All tensors on one HIP device (the product path has no CPU kernels)."""
    dev = None
    for t in tensors:
        if t is None:
            continue
        if t.device.type != "cuda":
            raise RuntimeError(
                "pygrog_amd ops run on the GPU only: got a tensor on "
                f"'{t.device}'. Move inputs to a HIP device (torch 'cuda')."
            )
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise RuntimeError(f"tensors on different devices: {dev} vs {t.device}")
    if dev is None:
        raise RuntimeError("no tensor argument")
    return dev


_hip_seen = False


def hip_available() -> bool:
    """torch.cuda.is_available(), remembered once it has been true (the
    device count of a process does not drop)."""
    global _hip_seen
    if not _hip_seen:
        _hip_seen = bool(torch.cuda.is_available())
    return _hip_seen


def hip_device(device=None) -> torch.device:
    """Resolve the compute device: a HIP device, or raise."""
    if not hip_available():
        raise RuntimeError(
            "pygrog_amd needs a HIP (ROCm) GPU: torch.cuda.is_available() is False"
        )
    if isinstance(device, torch.device) and device.type == "cuda" and device.index is not None:
        return device
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    device = torch.device(device)
    if device.type != "cuda":
        raise RuntimeError(f"pygrog_amd computes on HIP devices, not '{device}'")
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return device


def _check(cond: bool, *msg) -> None:
    # Same exception type as the reference's TORCH_CHECK.
    if not cond:
        raise RuntimeError("".join(str(m) for m in msg))


def _real_of(dtype: torch.dtype) -> torch.dtype:
    return torch.float32 if dtype == torch.complex64 else torch.float64


def _sfx(dtype: torch.dtype) -> str:
    if dtype == torch.complex64:
        return "c64"
    if dtype == torch.complex128:
        return "c128"
    raise RuntimeError(f"expected ComplexFloat or ComplexDouble, got {dtype}")


class TorchExt:
    """The reference's ``_pygrog_torch`` op set on HIP (module.cpp:54-87)."""

    # -- grog_interpolate (csrc/torch/grog_interp.cpp:112-149) --------------
    @staticmethod
    def grog_interpolate(data: torch.Tensor, indexes: torch.Tensor, kernel: torch.Tensor) -> torch.Tensor:
        _check(data.dim() == 3, "data must be 3-D (nsamples, nbatches, ncoils), got ", data.dim(), "-D")
        _check(indexes.dim() == 1, "indexes must be 1-D (nsamples,), got ", indexes.dim(), "-D")
        _check(kernel.dim() == 3, "kernel must be 3-D (n_kernels, ncoils, ncoils), got ", kernel.dim(), "-D")
        _check(indexes.dtype == torch.int64, "indexes must be int64, got ", indexes.dtype)
        _check(data.size(0) == indexes.size(0), "data.size(0) (", data.size(0), ") != indexes.size(0) (", indexes.size(0), ")")
        _check(
            data.size(2) == kernel.size(1) and kernel.size(1) == kernel.size(2),
            "ncoils mismatch between data (", data.size(2), ") and kernel (", kernel.size(1), "x", kernel.size(2), ")",
        )
        _check(kernel.dtype == data.dtype, "kernel dtype ", kernel.dtype, " != data dtype ", data.dtype)
        sfx = _sfx(data.dtype)
        dev = require_hip(data, indexes, kernel)
        data = data.contiguous()
        indexes = indexes.contiguous()
        kernel = kernel.contiguous()
        out = torch.empty_like(data)
        N, B, C = data.shape
        call(f"pg_grog_interpolate_{sfx}", ptr(data), ptr(indexes), ptr(kernel), ptr(out), N, B, C, kernel.size(0), stream(dev))
        return out

    # -- scatter / gather (csrc/torch/sparse_ops.cpp:233-340) ---------------
    @staticmethod
    def _scatter_checks(grid, data, indices, weights):
        _check(grid.dim() == 1, "grid must be 1-D, got ", grid.dim(), "-D")
        _check(data.dim() == 1, "data must be 1-D, got ", data.dim(), "-D")
        _check(indices.dim() == 1, "indices must be 1-D")
        _check(weights.dim() == 1, "weights must be 1-D")
        _check(data.size(0) == indices.size(0), "data/indices size mismatch")
        _check(data.size(0) == weights.size(0), "data/weights size mismatch")
        _check(indices.dtype == torch.int64, "indices must be int64")
        sfx = _sfx(data.dtype)
        _check(grid.dtype == data.dtype, "grid/data dtype mismatch")
        _check(weights.dtype == _real_of(data.dtype), "weights dtype must be ", _real_of(data.dtype), ", got ", weights.dtype)
        _check(grid.is_contiguous(), "grid must be contiguous (updated in place)")
        return sfx

    @staticmethod
    def scatter_add(grid, data, indices, weights) -> None:
        sfx = TorchExt._scatter_checks(grid, data, indices, weights)
        dev = require_hip(grid, data, indices, weights)
        data, indices, weights = data.contiguous(), indices.contiguous(), weights.contiguous()
        call(f"pg_scatter_add_{sfx}", ptr(grid), grid.numel(), ptr(data), ptr(indices), ptr(weights), data.numel(), stream(dev))

    @staticmethod
    def scatter_add_binned(grid, data, indices, weights, bin_starts, bin_size: int) -> None:
        sfx = TorchExt._scatter_checks(grid, data, indices, weights)
        _check(bin_starts.dim() == 1, "bin_starts must be 1-D")
        _check(bin_starts.dtype == torch.int64, "bin_starts must be int64")
        _check(bin_size > 0, "bin_size must be positive")
        dev = require_hip(grid, data, indices, weights, bin_starts)
        data, indices, weights = data.contiguous(), indices.contiguous(), weights.contiguous()
        bin_starts = bin_starts.contiguous()
        call(
            f"pg_scatter_add_binned_{sfx}", ptr(grid), grid.numel(), ptr(data), ptr(indices), ptr(weights),
            data.numel(), ptr(bin_starts), bin_starts.numel(), int(bin_size), stream(dev),
        )

    @staticmethod
    def gather(grid, indices, weights) -> torch.Tensor:
        _check(grid.dim() == 1, "grid must be 1-D, got ", grid.dim(), "-D")
        _check(indices.dim() == 1, "indices must be 1-D")
        _check(weights.dim() == 1, "weights must be 1-D")
        _check(indices.size(0) == weights.size(0), "indices/weights size mismatch")
        _check(indices.dtype == torch.int64, "indices must be int64")
        sfx = _sfx(grid.dtype)
        _check(weights.dtype == _real_of(grid.dtype), "weights dtype must be ", _real_of(grid.dtype), ", got ", weights.dtype)
        dev = require_hip(grid, indices, weights)
        grid, indices, weights = grid.contiguous(), indices.contiguous(), weights.contiguous()
        out = torch.zeros(indices.numel(), dtype=grid.dtype, device=dev)
        call(f"pg_gather_{sfx}", ptr(grid), grid.numel(), ptr(indices), ptr(weights), indices.numel(), ptr(out), stream(dev))
        return out

    # -- Toeplitz PSF builders (csrc/torch/toep_psf.cpp:242-367) ------------
    @staticmethod
    def psf_scatter_scalar(psf, indices, w_sq) -> None:
        _check(psf.dim() == 1, "psf must be 1-D, got ", psf.dim(), "-D")
        _check(indices.dim() == 1, "indices must be 1-D")
        _check(w_sq.dim() == 1, "w_sq must be 1-D")
        _check(indices.size(0) == w_sq.size(0), "indices/w_sq size mismatch")
        _check(indices.dtype == torch.int64, "indices must be int64")
        _check(psf.dtype == w_sq.dtype, "psf/w_sq dtype mismatch")
        _check(psf.dtype in (torch.float32, torch.float64), "psf must be float32 or float64")
        _check(psf.is_contiguous(), "psf must be contiguous (updated in place)")
        dev = require_hip(psf, indices, w_sq)
        sfx = "f32" if psf.dtype == torch.float32 else "f64"
        indices, w_sq = indices.contiguous(), w_sq.contiguous()
        call(f"pg_psf_scatter_scalar_{sfx}", ptr(psf), psf.numel(), ptr(indices), ptr(w_sq), indices.numel(), stream(dev))

    @staticmethod
    def psf_scatter_outer(psf, indices, w_sq, basis_per_sample) -> None:
        _check(psf.dim() == 3, "psf must be 3-D (grid_size, M, M)")
        _check(psf.size(1) == psf.size(2), "psf last two dims must be equal")
        _check(indices.dim() == 1 and w_sq.dim() == 1, "indices and w_sq must be 1-D")
        _check(basis_per_sample.dim() == 2, "basis must be 2-D (N, M)")
        _check(
            indices.size(0) == w_sq.size(0) and indices.size(0) == basis_per_sample.size(0),
            "indices/w_sq/basis sample-count mismatch",
        )
        _check(basis_per_sample.size(1) == psf.size(1), "basis M does not match psf M")
        _check(indices.dtype == torch.int64, "indices must be int64")
        sfx = _sfx(psf.dtype)
        _check(basis_per_sample.dtype == psf.dtype, "basis/psf dtype mismatch")
        _check(w_sq.dtype == _real_of(psf.dtype), "w_sq dtype must be ", _real_of(psf.dtype))
        _check(psf.is_contiguous(), "psf must be contiguous (updated in place)")
        dev = require_hip(psf, indices, w_sq, basis_per_sample)
        indices, w_sq, b = indices.contiguous(), w_sq.contiguous(), basis_per_sample.contiguous()
        call(f"pg_psf_scatter_outer_{sfx}", ptr(psf), psf.size(0), psf.size(1), ptr(indices), ptr(w_sq), ptr(b), indices.numel(), stream(dev))

    @staticmethod
    def psf_scatter_outer_basis(psf, indices, w_sq, basis, time_index) -> None:
        _check(psf.dim() == 3, "psf must be 3-D (grid_size, K, K)")
        _check(psf.size(1) == psf.size(2), "psf last two dims must be equal")
        _check(basis.dim() == 2, "basis must be 2-D (K, T)")
        _check(basis.size(0) == psf.size(1), "basis K does not match psf K")
        _check(indices.dim() == 1 and w_sq.dim() == 1 and time_index.dim() == 1, "indices, w_sq, time_index must be 1-D")
        _check(
            indices.size(0) == w_sq.size(0) and indices.size(0) == time_index.size(0),
            "indices/w_sq/time_index size mismatch",
        )
        _check(indices.dtype == torch.int64, "indices must be int64")
        _check(time_index.dtype == torch.int64, "time_index must be int64")
        sfx = _sfx(psf.dtype)
        _check(basis.dtype == psf.dtype, "basis/psf dtype mismatch")
        _check(w_sq.dtype == _real_of(psf.dtype), "w_sq dtype must be ", _real_of(psf.dtype))
        _check(psf.is_contiguous(), "psf must be contiguous (updated in place)")
        dev = require_hip(psf, indices, w_sq, basis, time_index)
        indices, w_sq, basis, time_index = (x.contiguous() for x in (indices, w_sq, basis, time_index))
        call(
            f"pg_psf_scatter_outer_basis_{sfx}", ptr(psf), psf.size(0), psf.size(1), ptr(indices), ptr(w_sq),
            ptr(basis), basis.size(1), ptr(time_index), indices.numel(), stream(dev),
        )


_ext = TorchExt()


def get_torch_ext():
    """The drop-in for the reference's ``_get_torch_ext()``: the seven ops as
    torch custom operators (``torch.ops.pygrog_amd``, registered on first
    use; see ``pygrog_amd._torch_ops``), after making sure the HIP library
    loads."""
    load()
    from . import _torch_ops

    return _torch_ops.ops()


# Workspace helper for the two-phase hipCUB entry points.
def workspace_bytes(name: str, *args) -> int:
    size = ctypes.c_size_t(0)
    call(name, None, ctypes.byref(size), *args)
    return int(size.value)


def run_with_workspace(name: str, device: torch.device, *args) -> None:
    nbytes = workspace_bytes(name, *args)
    ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
    size = ctypes.c_size_t(nbytes)
    call(name, ptr(ws), ctypes.byref(size), *args)
    # keep ws alive until the enqueued work has consumed it
    ws.record_stream(torch.cuda.current_stream(device))


def env_flag(name: str, default: str = "") -> str:
    return os.environ.get(name, default)
