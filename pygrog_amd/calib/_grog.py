"""GROG interpolator and plan (drop-in for ``pygrog.calib``), device-resident.

Public surface and semantics follow the reference src/pygrog/calib/_grog.py:
``GrogPlan`` (:27-181), ``GrogInterpolator`` (:270-912) with
``calc_interp_table`` (:382-422), ``interpolate`` (:437-555), shot-by-shot
gridding (:697-731, 904-912), and the plan construction of
``_create_plan`` / ``_create_plan_single`` / ``_attach_fft_plan``
(:918-1234).

Where the reference builds the plan with float32 torch ops on the host and
grids with per-coil Python loops over C++ scatter calls, here:
  * the plan (targets, shifts, EGROG weights, sort permutations, the per-cell
    CSR and the chunked reduction units) is compiled by HIP kernels
    (``pygrog_amd/csrc/plan.hip``) with the reference's float32 operation
    order, so target indices and shifts are bit-exact;
  * ``interpolate(grid=True)`` / ``ret_image=True`` run one fused kernel
    (operator apply + deterministic cell-ordered reduction) for all batch,
    stack and coil entries (``pygrog_amd/csrc/grid.hip``), or the BF16-MFMA
    grouped-GEMM path (``grid_mfma.hip``) with ``mode="bf16"``/``"bf16x3"``/
    ``"bf16x3c"`` (the last over block-floating-point sample and V rows);
  * ``interpolate()`` (sparse output) runs one fused apply kernel.

Public tensor fields of the plan (``target_idx``, ``indices``, ...) live on
the GPU and are copied to the plan's home device (the device of the input
coordinates, CPU for numpy input) on first access.
"""

from __future__ import annotations

__all__ = ["GrogInterpolator", "GrogPlan"]

import copy
import ctypes
import functools
import gc
import math
from typing import NamedTuple
from types import SimpleNamespace

import numpy as np
import torch

from .. import _lib
from ._grappa import build_table, quant_steps

# Replicas folded sequentially per reduction unit; fixes the summation order
# of every cell independently of tiling (see grid.hip).
UNIT_CHUNK = 32

# Largest shift a replica may have along an axis (grid steps); the plan of
# every interpolator carries it (reference _grog.py:1123-1124).
PLAN_RADIUS = 0.5

_LAZY_FIELDS = ("target_idx", "distances", "weights", "indices", "sqrt_weights", "sort_perm", "inv_perm", "time_map")


# ---------------------------------------------------------------------------
# Plan container
# ---------------------------------------------------------------------------
class GrogPlan(SimpleNamespace):
    """GROG / FFT plan (reference ``GrogPlan``, _grog.py:27-181).

    Scalar fields (``grid_shape``, ``image_shape``, ``n_samples``, ...) are
    plain attributes.  Trajectory-bound tensors are held on the GPU in
    ``_dev`` and materialised on the home device on first attribute access.
    """

    def __getattr__(self, name):
        d = self.__dict__
        if name in _LAZY_FIELDS:
            dev = d.get("_dev")
            if dev is not None and name in dev and dev[name] is not None:
                t = dev[name]
                home = d.get("_home")
                v = t if (home is None or t.device == home) else t.to(home)
                d[name] = v
                return v
            if name == "time_map":
                return None
        raise AttributeError(name)

    @property
    def stack_shape(self) -> tuple:
        return tuple(getattr(self, "_stack_shape", ()))

    @stack_shape.setter
    def stack_shape(self, value):
        self._stack_shape = tuple(int(s) for s in (value or ()))

    @property
    def is_stacked(self) -> bool:
        return len(self.stack_shape) > 0

    @property
    def pre_weights(self) -> torch.Tensor:
        """``sqrt_weights`` in interpolate-output order: ``sqrt_w[inv_perm]``
        (reference _grog.py:60-76)."""
        sw = self.sqrt_weights
        ip = self.inv_perm
        if not self.is_stacked:
            return sw[ip]
        return torch.gather(sw, -1, ip)

    @classmethod
    def _stack(cls, plans: list) -> "GrogPlan":
        """Stack unstacked per-element plans (reference _grog.py:80-181)."""
        if not plans:
            raise ValueError("Cannot stack an empty list of plans.")
        for p in plans:
            if getattr(p, "stack_shape", ()) != ():
                raise ValueError("GrogPlan.stack requires unstacked input plans.")
        ref = plans[0]
        global_fields = ("shape", "oversamp", "grid_shape", "kernel_width", "kernel_shape", "radius",
                         "image_shape", "grid_size", "n_samples", "natural_shape")
        for f in global_fields:
            ref_v = getattr(ref, f, None)
            for p in plans[1:]:
                if getattr(p, f, None) != ref_v:
                    raise ValueError(f"GrogPlan.stack: field '{f}' differs across plans")
        out = cls(**{f: getattr(ref, f) for f in global_fields})
        out.stack_shape = (len(plans),)
        home = ref.__dict__.get("_home")
        out._home = home

        def dev_or_attr(p, name):
            d = p.__dict__.get("_dev") or {}
            t = d.get(name)
            return t if t is not None else torch.as_tensor(getattr(p, name))

        dev = {}
        for name in ("target_idx", "weights", "indices", "sqrt_weights", "sort_perm", "inv_perm"):
            dev[name] = torch.stack([dev_or_attr(p, name) for p in plans], dim=0)
        if all(getattr(p, "distances", None) is not None for p in plans):
            dev["distances"] = torch.stack([dev_or_attr(p, "distances") for p in plans], dim=0)
        else:
            out.distances = None
        tms = [getattr(p, "time_map", None) for p in plans]
        if any(t is not None for t in tms):
            if any(t is None for t in tms):
                raise ValueError("GrogPlan.stack: all plans must agree on time_map presence")
            dev["time_map"] = torch.stack([torch.as_tensor(t) for t in tms], dim=0)
        out._dev = dev
        return out


# ---------------------------------------------------------------------------
# Helpers (reference _grog.py:1240-1322)
# ---------------------------------------------------------------------------
def _kernel_offsets(ndim: int, width: int, shape: str) -> np.ndarray:
    """Integer neighbourhood offsets, (kw, ndim) int32 (reference :1240-1277).

    Axis values ``arange(width) - floor((width - 1) / 2)``, enumerated with the
    last axis fastest; 'circle'/'sphere' keeps radius <= (width-1)/2 + 0.5,
    'cross' keeps offsets with at most one nonzero axis.  The array is
    shared between calls with the same arguments (read-only).
    """
    return _kernel_offsets_of(ndim, width, shape)


@functools.lru_cache(maxsize=64)
def _kernel_offsets_of(ndim: int, width: int, shape: str) -> np.ndarray:
    offs = _kernel_offsets_new(ndim, width, shape)
    offs.flags.writeable = False
    return offs


def _kernel_offsets_new(ndim: int, width: int, shape: str) -> np.ndarray:
    half = (width - 1) / 2.0
    ax = np.arange(width) - int(np.floor(half))
    mesh = np.meshgrid(*([ax] * ndim), indexing="ij")
    offs = np.stack([m.ravel() for m in mesh], axis=-1).astype(np.int32)
    key = shape.lower()
    if key in ("square", "cube"):
        return offs
    if key in ("circle", "sphere"):
        r = np.sqrt((offs.astype(np.float32) ** 2).sum(axis=-1))
        return offs[r <= half + 0.5]
    if key == "cross":
        return offs[(offs != 0).sum(axis=-1) <= 1]
    raise ValueError(
        f"Unknown kernel_shape '{shape}'. Use 'circle'/'sphere', 'square'/'cube', or 'cross'."
    )


def _default_shape(ndim, shape):
    if np.isscalar(shape):
        shape = ndim * [shape]
    return tuple(int(s) for s in shape)


def _default_oversamp(ndim, oversamp):
    if oversamp is None:
        oversamp = (1.0, 1.0) if ndim == 2 else (1.0, 1.0, 1.2)
    if np.isscalar(oversamp):
        oversamp = ndim * [oversamp]
    oversamp = tuple(oversamp)
    if len(oversamp) != ndim:
        raise ValueError(f"Oversampling {oversamp} does not match ndim={ndim}")
    return oversamp


def _u32(t: torch.Tensor) -> int:
    return _lib.ptr(t)


def _exclusive_scan(counts: torch.Tensor, device) -> torch.Tensor:
    """counts: (n + 1,) int32 with counts[n] == 0 -> (n + 1,) exclusive sums."""
    n = counts.numel() - 1
    out = torch.empty_like(counts)
    _lib.run_with_workspace("pg_exclusive_scan_u32", device, _u32(counts), _u32(out), n, _lib.stream(device))
    return out


def compile_units(cell_ptr: torch.Tensor, S: int, G: int, device, chunk: int = UNIT_CHUNK,
                  unit_count: torch.Tensor | None = None):
    """Chunked reduction units from a per-element CSR (S, G + 1) uint32.

    Returns (cell_unit (S*G + 1), unit_cell (U), unit_rep (U + S)) int32
    tensors holding uint32 values; unit u of element s spans replicas
    [unit_rep[u + s], unit_rep[u + s + 1]).
    """
    st = _lib.stream(device)
    if unit_count is None:
        unit_count = torch.zeros(S * G + 1, dtype=torch.int32, device=device)
        _lib.call("pg_unit_count", _u32(cell_ptr), S, G, chunk, _u32(unit_count), st)
    cell_unit = _exclusive_scan(unit_count, device)
    U = int(cell_unit[-1].item())
    unit_cell = torch.empty(max(U, 1), dtype=torch.int32, device=device)
    unit_rep = torch.empty(U + S, dtype=torch.int32, device=device)
    _lib.call("pg_plan_units", _u32(cell_ptr), _u32(cell_unit), S, G, chunk, _u32(unit_cell), _u32(unit_rep), st)
    return cell_unit, unit_cell, unit_rep


# Cells with at least this many reduction units are "hot": reduced by the
# pre-pass kernels instead of inside their tile.
HOT_MIN_UNITS = 2

# arrays of one compiled plan's workspace (pg_plan_compile_layout)
PLAN_FIELDS = 19
_ITEMSIZE = {torch.int64: 8, torch.int32: 4, torch.float32: 4}


def hot_lists(cell_unit: torch.Tensor, S: int, G: int, device, min_units: int | None = None, U: int | None = None):
    """Hot-cell work lists per stack element.

    Returns (unit_hot (U) int32: element-local hot-cell index of each unit or
    -1, and per element a dict with hot_units (HU) int32 global unit ids in
    cell then unit order, hot_ptr (H + 1) int32, n_hot, n_hot_units).
    """
    if min_units is None:
        min_units = HOT_MIN_UNITS
    cu = cell_unit.to(torch.int64)
    if U is None:
        U = int(cu[-1].item())
    cnt = cu[1:] - cu[:-1]
    unit_hot = torch.full((max(U, 1),), -1, dtype=torch.int32, device=device)
    per = []
    hot_all = torch.nonzero(cnt >= min_units).squeeze(1)
    for s in range(S):
        hs = hot_all[(hot_all >= s * G) & (hot_all < (s + 1) * G)]
        lens = cnt[hs]
        H = int(hs.numel())
        ptr = torch.zeros(H + 1, dtype=torch.int64, device=device)
        if H:
            ptr[1:] = torch.cumsum(lens, 0)
        HU = int(ptr[-1].item()) if H else 0
        if HU:
            owner = torch.repeat_interleave(torch.arange(H, device=device), lens, output_size=HU)
            units = cu[hs][owner] + (torch.arange(HU, device=device) - ptr[:-1][owner])
            unit_hot[units] = owner.to(torch.int32)
            units = units.to(torch.int32)
        else:
            units = torch.zeros(1, dtype=torch.int32, device=device)
        per.append({"hot_units": units.contiguous(), "hot_ptr": ptr.to(torch.int32).contiguous(), "n_hot": H,
                    "n_hot_units": HU})
    return unit_hot, per


def ordinary_rows(cell_ptr: torch.Tensor, cell_unit: torch.Tensor, S: int, G: int,
                  min_units: int | None = None) -> torch.Tensor:
    """(S, G + 1) int32: replicas of ordinary cells (one unit, not hot)
    before each cell, per stack element -- the tile kernels' row numbering,
    which leaves out the cells the hot pre-pass or the multi-unit path
    reduce.  ``min_units`` must be the one ``hot_lists`` used."""
    if min_units is None:
        min_units = HOT_MIN_UNITS
    cp = cell_ptr.reshape(S, G + 1).to(torch.int64)
    cnt = cp[:, 1:] - cp[:, :-1]
    cu = cell_unit.to(torch.int64)
    nu = (cu[1:] - cu[:-1]).reshape(S, G)
    special = (nu >= max(int(min_units), 1)) | (nu > 1)
    vc = torch.where(special, torch.zeros_like(cnt), cnt)
    out = torch.zeros(S, G + 1, dtype=torch.int64, device=cell_ptr.device)
    out[:, 1:] = torch.cumsum(vc, 1)
    return out.to(torch.int32).contiguous()


class KernelTimer:
    """Per-kernel HIP-event timing on the launch stream: ``mark(name)``
    brackets one launch; ``durations()`` (after a synchronize) returns the
    milliseconds of every bracketed launch, per kernel name."""

    def __init__(self, device):
        self.device = device
        self.events: list = []

    def start(self):
        e = torch.cuda.Event(enable_timing=True)
        e.record(torch.cuda.current_stream(self.device))
        return e

    def stop(self, name, e0):
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record(torch.cuda.current_stream(self.device))
        self.events.append((name, e0, e1))

    def durations(self) -> dict:
        out: dict = {}
        for name, a, b in self.events:
            out.setdefault(name, []).append(a.elapsed_time(b))
        return out

    def reset(self):
        self.events = []


class EntryGroups(NamedTuple):
    """Replicas grouped by table entry, for the grouped MFMA apply."""

    ent_ptr: torch.Tensor   # (K + 1) int32: group of entry e = positions [ent_ptr[e], ent_ptr[e+1])
    ent_rep: torch.Tensor   # (R) int32: replica (cell-sorted index) at each position
    items: torch.Tensor     # (>= n_items, 2) int32: (entry, first offset) work chunks, the first n_items used
    n_items: int
    ent_src: torch.Tensor   # (R) int32: source sample of each position's replica
    ent_s: torch.Tensor     # (R) float32: sqrt weight of each position's replica


def group_by_entry(rep_ent: torch.Tensor, R: int, K: int, device, chunk: int, rep_src: torch.Tensor,
                   rep_s: torch.Tensor, defer: bool = False) -> EntryGroups:
    """Replicas of each table entry (a stable radix sort of ``rep_ent[:R]``),
    the work list of chunks of ``chunk`` replicas, and each replica's source
    and weight gathered into that order (so the apply reads its records
    contiguously instead of gathering them): one library call
    (``pg_group_by_entry``), one small host read.  ``defer``: no host read
    (``pg_group_by_entry_async``): ``n_items`` is the work list's capacity,
    whose unused slots the apply skips."""
    if R <= 0:
        z = torch.zeros(1, dtype=torch.int32, device=device)
        return EntryGroups(torch.zeros(K + 1, dtype=torch.int32, device=device), z, torch.zeros(1, 2, dtype=torch.int32,
                           device=device), 0, z.clone(), torch.zeros(1, dtype=torch.float32, device=device))
    keys, src, sw = rep_ent[:R], rep_src[:R], rep_s[:R]
    if keys.dtype != torch.int32 or not keys.is_contiguous():
        keys = keys.to(torch.int32).contiguous()
    if src.dtype != torch.int32 or not src.is_contiguous():
        src = src.to(torch.int32).contiguous()
    if sw.dtype != torch.float32 or not sw.is_contiguous():
        sw = sw.to(torch.float32).contiguous()
    im, tb = ctypes.c_int64(0), ctypes.c_int64(0)
    _lib.call("pg_group_by_entry_layout", R, K, chunk, ctypes.addressof(im), ctypes.addressof(tb))
    ent_ptr = torch.empty(K + 1, dtype=torch.int32, device=device)
    ent_rep = torch.empty(R, dtype=torch.int32, device=device)
    ent_src = torch.empty(R, dtype=torch.int32, device=device)
    ent_s = torch.empty(R, dtype=torch.float32, device=device)
    items = torch.empty(int(im.value), 2, dtype=torch.int32, device=device)
    tmp = torch.empty(int(tb.value), dtype=torch.uint8, device=device)
    args = (_lib.ptr(keys), _lib.ptr(src), _lib.ptr(sw), R, K, chunk, _lib.ptr(ent_ptr), _lib.ptr(ent_rep),
            _lib.ptr(ent_src), _lib.ptr(ent_s), _lib.ptr(items), int(im.value), _lib.ptr(tmp), int(tb.value))
    if defer:
        _lib.call("pg_group_by_entry_async", *args, None, _lib.stream(device))
        return EntryGroups(ent_ptr, ent_rep, items, int(im.value), ent_src, ent_s)
    n = ctypes.c_int64(0)
    _lib.call("pg_group_by_entry", *args, ctypes.addressof(n), _lib.stream(device))
    return EntryGroups(ent_ptr, ent_rep, items, int(n.value), ent_src, ent_s)


# The BF16 modes can keep their sample-major copy of the k-space in the
# library's 24-bit row format (fp32 rounded to nearest even at 16 significant
# bits, 3 bytes a value, rows on 16-byte boundaries): the split products use
# 16 bits of an input anyway, and the transpose writes 3/4 of the bytes.  The
# grouped apply then reads one row per replica at random, and such a read
# costs the whole 128-byte L2 lines the row touches (measured on MI355X: at 32
# coils 192-byte rows fetch as much as 256-byte ones), so the format is used
# where its rows touch no more lines on average than complex64 rows (32
# coils: 2 and 2; 16 coils: 1.5 against 1, and the apply was slower), and up
# to this many coils (above, the 24-bit apply kernel runs out of registers).
S24_MAX_COILS = 32


@functools.lru_cache(maxsize=None)
def _lines_per_row(row_bytes: int) -> float:
    """128-byte lines a row of ``row_bytes`` touches, averaged over the
    offsets rows at a stride of ``row_bytes`` take."""
    offs = sorted({(k * row_bytes) % 128 for k in range(128)})
    return sum((o + row_bytes + 127) // 128 for o in offs) / len(offs)


def uses_s24(mode: str, C: int) -> bool:
    """Whether gridding mode ``mode`` at ``C`` coils uses 24-bit sample rows."""
    if mode in ("fp32", "bf16x3c") or C > S24_MAX_COILS:
        return False
    return _lines_per_row(16 * ((6 * C + 15) // 16)) <= _lines_per_row(8 * C)


# Gridding modes: the plain fused FP32 kernel, and the grouped BF16-MFMA
# path with one product (bf16), three split products (bf16x3) over rows
# holding 16 significant bits of every value (24-bit or complex64 rows), or
# three split products over block-floating-point rows (bf16x3c: 15-bit
# mantissas with an exponent per 4 coils, 16 * ceil(C / 4) bytes a row --
# one 128-byte line at 32 coils; each value within 2^-14 of its 4-coil
# group's largest magnitude).
GRID_MODES = ("fp32", "bf16", "bf16x3", "bf16x3c")


def mode_terms(mode: str) -> int:
    """Split-BF16 products of a grouped mode."""
    return 1 if mode == "bf16" else 3


def row_format(mode: str, C: int) -> str:
    """Sample-row format of ``mode`` at ``C`` coils: "c64", "s24" or "bfp"."""
    if mode == "bf16x3c":
        return "bfp"
    return "s24" if uses_s24(mode, C) else "c64"


_ROW_DTYPE = {"c64": torch.complex64, "s24": torch.uint8, "bfp": torch.int16}
_ROW_TRANSPOSE = {"c64": "_c64", "s24": "_s24", "bfp": "_bfp"}
_ROW_APPLY = {"c64": "", "s24": "_s24", "bfp": "_bfp"}


def row_format_of(yt: torch.Tensor) -> str:
    """Format of a sample-row buffer made by ``new_sample_rows``."""
    return {v: k for k, v in _ROW_DTYPE.items()}[yt.dtype]


def transpose_entry(fmt: str, frames: bool) -> str:
    return ("pg_transpose_frames" if frames else "pg_transpose_cn") + _ROW_TRANSPOSE[fmt]


def apply_entry(fmt: str) -> str:
    return "pg_apply_grouped_bf16" + _ROW_APPLY[fmt]


@functools.lru_cache(maxsize=None)
def sample_row_bytes(C: int, fmt: str) -> int:
    """Bytes of one sample's row of the sample-major copy."""
    if fmt == "s24":
        return int(_lib.load().pg_s24_row_bytes(C))
    if fmt == "bfp":
        return int(_lib.load().pg_bfp_row_bytes(C))
    return 8 * C


def v_format(mode: str) -> str:
    """Row format of a grouped mode's applied-replica rows V: bf16x3c keeps
    them in the block-floating-point format too (half the bytes of
    complex64 rows both ways: written by the apply, read by the reduction),
    the other modes in complex64."""
    return "bfp" if mode == "bf16x3c" else "c64"


def reduce_entry(mode: str) -> str:
    """The reduction (or, for fp32, the fused gridding) entry of ``mode``."""
    if mode == "fp32":
        return "pg_grid_fp32"
    return "pg_grid_reduce_v_bfp" if v_format(mode) == "bfp" else "pg_grid_reduce_v"


# cells per wave-tile launch (four wavefronts of 64 cells)
WAVE_TILE = 256


def mode_tile_options(mode: str, C: int) -> list:
    """Launch tiles of ``mode``'s gridding kernel: block-floating-point V
    rows are reduced by the wave-tile kernels only."""
    return list(_mode_tiles(mode, int(C))[0])


def mode_default_tile(mode: str, C: int) -> int:
    return _mode_tiles(mode, int(C))[1]


@functools.lru_cache(maxsize=None)
def _mode_tiles(mode, C: int) -> tuple:
    """(tile options, default tile) of a mode and coil count (library
    constants: asked once)."""
    opts = _lib.tile_options(C)
    if mode is not None and v_format(mode) == "bfp":
        opts = [t for t in opts if t == WAVE_TILE]
    t = int(_lib.load().pg_grid_tile_cells(C))
    return tuple(opts), (t if t in opts else opts[0])



def new_sample_rows(n: int, C: int, fmt: str, device) -> torch.Tensor:
    """Buffer for the sample-major copy of ``n`` samples in format ``fmt``."""
    dt = _ROW_DTYPE[fmt]
    per = sample_row_bytes(C, fmt) // torch.empty(0, dtype=dt).element_size()
    return torch.empty(max(n, 1) * per, dtype=dt, device=device)


def _to_device_async(t: torch.Tensor, device) -> torch.Tensor:
    """Copy a small host tensor to ``device`` on the current stream without
    waiting (through page-locked memory, which torch keeps until the copy
    has run)."""
    if not _lib.hip_available():
        return t.to(device)
    return t.pin_memory().to(device, non_blocking=True)


class GridProgram:
    """A ready-to-launch fused gridding job: device frame table + tile map.

    Frames are (batch b, stack element s) pairs of one compiled plan; built
    once per (input, output) buffer pair so repeated launches only enqueue
    the kernel (no host work beyond the ctypes call).
    """

    def __init__(self, frames: list, C: int, G: int, device, tile_cells: int, n_blocks: int | None = None,
                 tile_opts: list | None = None):
        self.device = device
        self.C = int(C)
        self.tile_opts = list(tile_opts) if tile_opts is not None else _lib.tile_options(self.C)
        self.F = len(frames)
        self.tile_cells = int(tile_cells)
        # per-frame scratch of the hot-cell pre-pass
        self._scratch = []
        self.max_hot_units = 0
        self.max_hot = 0
        for f in frames:
            if f.n_hot_units > 0:
                pu = torch.empty(f.n_hot_units * self.C, dtype=torch.complex64, device=device)
                self._scratch.append(pu)
                f.p_unit = pu.data_ptr()
                self.max_hot_units = max(self.max_hot_units, int(f.n_hot_units))
                self.max_hot = max(self.max_hot, int(f.n_hot))
        self.G = int(G)
        # the frame table and the default tile map go up in one copy from
        # page-locked memory, without a blocking copy: a pageable copy would
        # wait for everything already on the stream (the serving loop builds
        # one program per batch while the previous batch runs)
        fb = bytes((_lib.Frame * self.F)(*frames))
        nf = (len(fb) + 15) // 16 * 16
        tp = self._tile_map(self.tile_cells)
        host = torch.empty(nf + tp.nbytes, dtype=torch.uint8, pin_memory=_lib.hip_available())
        hn = host.numpy()
        hn[:len(fb)] = np.frombuffer(fb, dtype=np.uint8)
        hn[nf:] = tp.view(np.uint8)
        both = host.to(device, non_blocking=True)
        self.frames = both[:nf]
        self._maps = {self.tile_cells: both[nf:].view(torch.int32)}
        self.set_shape(tile_cells, n_blocks)

    def _tile_map(self, tile_cells: int) -> np.ndarray:
        """First tile of every frame, (F + 1) int32."""
        tiles_per = (self.G + tile_cells - 1) // tile_cells
        return (np.arange(self.F + 1, dtype=np.int64) * tiles_per).astype(np.int32)

    def set_shape(self, tile_cells: int, n_blocks: int | None = None) -> None:
        """Launch shape: cells per tile (one of ``_lib.tile_options(C)``) and
        the persistent grid size (None: one workgroup per tile up to 12 per
        CU).  The shape never changes the result bits."""
        if int(tile_cells) not in self.tile_opts:
            raise ValueError(f"tile_cells {tile_cells} not available for C={self.C}: {self.tile_opts}")
        self.tile_cells = int(tile_cells)
        tiles_per = (self.G + self.tile_cells - 1) // self.tile_cells
        self.n_tiles = tiles_per * self.F
        # every tile map stays alive with the program (a graph captured
        # earlier may name any of them)
        if self.tile_cells not in self._maps:
            self._maps[self.tile_cells] = _to_device_async(torch.from_numpy(self._tile_map(self.tile_cells)),
                                                           self.device)
        self.tile_ptr = self._maps[self.tile_cells]
        self._mark_uploaded()
        if n_blocks is None:
            n_blocks = min(self.n_tiles, 256 * 12)
        self.n_blocks = max(1, int(n_blocks))

    def _mark_uploaded(self) -> None:
        """Record where the tables' uploads were enqueued; the next launch
        from another stream waits for them."""
        self._upload = None
        if _lib.hip_available():
            self._upload = (torch.cuda.current_stream(self.device), torch.cuda.Event())
            self._upload[1].record(self._upload[0])

    def launch(self, entry: str = "pg_grid_fp32", timer: "KernelTimer | None" = None) -> None:
        if self._upload is not None:
            up_stream, up_event = self._upload
            cur = torch.cuda.current_stream(self.device)
            if cur != up_stream:
                cur.wait_event(up_event)
            self._upload = None
        e0 = timer.start() if timer is not None else None
        _lib.call(entry, _lib.ptr(self.frames), _lib.ptr(self.tile_ptr), self.F, self.n_tiles, self.C,
                  self.tile_cells, self.n_blocks, self.max_hot_units, self.max_hot, _lib.stream(self.device))
        if timer is not None:
            timer.stop(entry, e0)


def tune_program(prog: GridProgram, entry: str, reps: int = 5, candidates=None) -> dict:
    """Time ``prog`` over its compiled tile sizes and candidate persistent
    grid sizes (the inputs the launch reads must already hold real data) and
    leave it at the fastest shape.  The shape never changes the bits."""
    n_cu = torch.cuda.get_device_properties(prog.device).multi_processor_count
    times = {}
    for tc in prog.tile_opts:
        prog.set_shape(tc, None)
        cands = candidates or [n_cu * k for k in (4, 8, 16, 32, 64)] + [prog.n_tiles]
        for nb in sorted({max(8, min(int(c), prog.n_tiles)) for c in cands}):
            prog.set_shape(tc, nb)
            prog.launch(entry)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                prog.launch(entry)
            e1.record()
            torch.cuda.synchronize(prog.device)
            times[(tc, nb)] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    prog.set_shape(*best)
    return {"tile_cells": best[0], "n_blocks": best[1], "ms": times[best],
            "tried": {f"{k[0]}x{k[1]}": v for k, v in times.items()}}


# ---------------------------------------------------------------------------
# Plan creation on the GPU
# ---------------------------------------------------------------------------
@functools.lru_cache(maxsize=256)
def _geom_of(shape: tuple, grid_shape: tuple, ndim: int, kernel_width, kernel_shape: str) -> _lib.PlanGeom:
    """The plan compiler's geometry argument (passed by value; shared
    between calls with the same arguments, never modified)."""
    return _geom(shape, grid_shape, _kernel_offsets(ndim, kernel_width, kernel_shape))


def _geom(shape, grid_shape, offsets) -> _lib.PlanGeom:
    ndim = len(shape)
    kw = offsets.shape[0]
    if kw > 128:
        raise ValueError(f"kernel has {kw} offsets; at most 128 are supported")
    g = _lib.PlanGeom()
    g.ndim = ndim
    g.kw = kw
    for d in range(3):
        if d < ndim:
            s, gs = shape[d], grid_shape[d]
            g.grid_shape[d] = gs
            g.strides[d] = int(np.prod(grid_shape[d + 1:])) if d + 1 < ndim else 1
            amp = 2 * (s - 1 - s // 2)
            g.half_amp[d] = float(np.float32(0.5) * np.float32(amp))
            g.origins[d] = float(np.float32(-(s // 2)))
            g.steps[d] = float(np.float32((s - 1) / (gs - 1) if gs > 1 else 1.0))
        else:
            g.grid_shape[d] = 1
            g.strides[d] = 0
            g.half_amp[d] = 0.0
            g.origins[d] = 0.0
            g.steps[d] = 1.0
    for k in range(kw):
        for d in range(ndim):
            g.offsets[k][d] = int(offsets[k, d])
    return g


def _create_plan(shape, coords, oversamp, kernel_width, kernel_shape, time_map, stack_shape=(), device=None):
    """Build the plan of every stack element in one set of kernel launches
    (reference _create_plan + _create_plan_single + _attach_fft_plan)."""
    return _plan_finish(_plan_launch(shape, coords, oversamp, kernel_width, kernel_shape, time_map, stack_shape,
                                     device))


class _PendingPlan(SimpleNamespace):
    """A plan whose kernels are enqueued but whose host summary may not have
    arrived yet (``_plan_launch(..., defer=True)``)."""

    def wait(self) -> None:
        if self.event is not None:
            self.event.synchronize()
            self.event = None


def _plan_launch(shape, coords, oversamp, kernel_width, kernel_shape, time_map, stack_shape=(), device=None,
                 defer: bool = False) -> _PendingPlan:
    """Enqueue the plan compiler on the current stream.  ``defer``: do not
    wait for it -- the summary goes to page-locked memory behind an event,
    and ``_plan_finish`` waits (the caller does host work meanwhile)."""
    dev = _lib.hip_device(device)
    st = _lib.stream(dev)
    home = coords.device if isinstance(coords, torch.Tensor) else torch.device("cpu")
    c = torch.as_tensor(coords)
    if c.dtype != torch.float32:
        c = c.to(torch.float32)
    ndim = int(c.shape[-1])
    s_shape = tuple(int(s) for s in (stack_shape or ()))
    if s_shape and tuple(c.shape[: len(s_shape)]) != s_shape:
        raise ValueError(f"coords leading shape {tuple(c.shape[:len(s_shape)])} does not match stack_shape {s_shape}")
    spatial = tuple(int(s) for s in c.shape[len(s_shape):-1])
    S = math.prod(s_shape)
    npts = math.prod(spatial)

    shape = _default_shape(ndim, shape)
    oversamp = _default_oversamp(ndim, oversamp)
    grid_shape = tuple(int(math.ceil(s * o)) for s, o in zip(shape, oversamp))
    G = math.prod(grid_shape)
    kw = int(_kernel_offsets(ndim, kernel_width, kernel_shape).shape[0])
    nper = npts * kw
    geom = _geom_of(shape, grid_shape, ndim, kernel_width, kernel_shape)

    cd = c.reshape(S, npts, ndim)
    if cd.device != dev:
        cd = cd.to(dev)
    cd = cd.contiguous()
    # the whole plan compiler in one library call (the calling thread holds
    # no interpreter lock meanwhile); every array is a view of one workspace
    offs = (ctypes.c_int64 * PLAN_FIELDS)()
    ob, tb = ctypes.c_int64(0), ctypes.c_int64(0)
    _lib.call("pg_plan_compile_layout", S, npts, ndim, kw, G, ctypes.addressof(offs), ctypes.addressof(ob),
              ctypes.addressof(tb))
    ws = torch.empty(int(ob.value), dtype=torch.uint8, device=dev)
    tmp = torch.empty(int(tb.value), dtype=torch.uint8, device=dev)
    distances = torch.empty(S, nper, ndim, dtype=torch.float32, device=dev)
    event = None
    if defer:
        summary_t = torch.empty(3 * S + 3, dtype=torch.int64, pin_memory=True)
        _lib.call("pg_plan_compile_async", _lib.ptr(cd), S, npts, geom, G, UNIT_CHUNK, int(HOT_MIN_UNITS),
                  _lib.ptr(ws), int(ob.value), _lib.ptr(tmp), int(tb.value), _lib.ptr(distances),
                  summary_t.data_ptr(), st)
        event = torch.cuda.Event()
        event.record(torch.cuda.current_stream(dev))
        summary = summary_t.numpy()
    else:
        summary_t = None
        summary = np.zeros(3 * S + 3, dtype=np.int64)
        _lib.call("pg_plan_compile", _lib.ptr(cd), S, npts, geom, G, UNIT_CHUNK, int(HOT_MIN_UNITS), _lib.ptr(ws),
                  int(ob.value), _lib.ptr(tmp), int(tb.value), _lib.ptr(distances), summary.ctypes.data, st)
    return _PendingPlan(dev=dev, S=S, npts=npts, ndim=ndim, kw=kw, nper=nper, G=G, s_shape=s_shape, spatial=spatial,
                        shape=shape, oversamp=oversamp, grid_shape=grid_shape, kernel_width=kernel_width,
                        kernel_shape=kernel_shape, time_map=time_map, home=home, offs=offs, ws=ws, tmp=tmp,
                        distances=distances, summary=summary, summary_t=summary_t, event=event, coords_dev=cd)


def _plan_finish(p: _PendingPlan):
    """The plan object of a launched plan compiler (waits for its summary)."""
    p.wait()
    p.tmp = None
    dev, S, npts, ndim, kw, nper, G = p.dev, p.S, p.npts, p.ndim, p.kw, p.nper, p.G
    s_shape, spatial, offs, ws, distances, summary = p.s_shape, p.spatial, p.offs, p.ws, p.distances, p.summary
    shape, oversamp, grid_shape, time_map = p.shape, p.oversamp, p.grid_shape, p.time_map
    kernel_width, kernel_shape, home = p.kernel_width, p.kernel_shape, p.home
    n_inv = [int(v) for v in summary[:S]]
    U = int(summary[S])
    hb = [int(v) for v in summary[S + 1:2 * S + 2]]
    ub = [int(v) for v in summary[2 * S + 2:3 * S + 3]]

    # one view of the workspace per element type (every field starts on a
    # 256-byte boundary), then a slice per field
    views = {}

    def field(i, dtype, *shape):
        n = 1
        for v in shape:
            n *= int(v)
        isz = _ITEMSIZE[dtype]
        start = int(offs[i]) // isz
        base = views.get(dtype)
        if base is None:
            base = views[dtype] = ws.view(dtype)
        return base[start:start + n].view(shape)

    i32 = torch.int32
    target_idx = field(0, torch.int64, S, nper)
    weights = field(1, torch.float32, S, nper)
    indices = field(2, torch.int64, S, nper)
    sqrt_weights = field(3, torch.float32, S, nper)
    sort_perm = field(4, torch.int64, S, nper)
    inv_perm = field(5, torch.int64, S, nper)
    cell_ptr = field(6, i32, S, G + 1)
    rep_src = field(7, i32, S, max(nper, 1))
    rep_s = field(8, torch.float32, S, max(nper, 1))
    density = field(9, torch.float32, S, G)
    cell_unit = field(10, i32, S * G + 1)
    unit_cell = field(11, i32, max(U, 1))
    unit_rep = field(12, i32, U + S)
    unit_hot = field(13, i32, max(U, 1))
    hot_units_all = field(14, i32, max(ub[S], 1))
    hot_ptr_all = field(15, i32, hb[S] + S)
    cell_vrow = field(16, i32, S, G + 1)
    n_invalid = field(17, torch.int64, S)
    hot = []
    for s in range(S):
        H, HU = hb[s + 1] - hb[s], ub[s + 1] - ub[s]
        units = hot_units_all[ub[s]:ub[s] + HU] if HU else torch.zeros(1, dtype=i32, device=dev)
        hot.append({"hot_units": units, "hot_ptr": hot_ptr_all[hb[s] + s:hb[s] + s + H + 1], "n_hot": H,
                    "n_hot_units": HU})

    if time_map is not None:
        tm = torch.as_tensor(np.asarray(time_map, dtype=np.float32) if not isinstance(time_map, torch.Tensor) else time_map)
        tm = tm.to(torch.float32).reshape(S, npts, 1).expand(S, npts, kw).reshape(S, nper).to(dev).contiguous()
    else:
        tm = None

    full = tuple(s_shape) + spatial + (kw,)
    dev_fields = {
        "target_idx": target_idx.reshape(full),
        "distances": distances.reshape(full + (ndim,)),
        "weights": weights.reshape(full),
        "time_map": tm.reshape(full) if tm is not None else None,
        "indices": indices.reshape(s_shape + (nper,)) if s_shape else indices.reshape(nper),
        "sqrt_weights": sqrt_weights.reshape(s_shape + (nper,)) if s_shape else sqrt_weights.reshape(nper),
        "sort_perm": sort_perm.reshape(s_shape + (nper,)) if s_shape else sort_perm.reshape(nper),
        "inv_perm": inv_perm.reshape(s_shape + (nper,)) if s_shape else inv_perm.reshape(nper),
    }
    plan = GrogPlan(shape=shape, oversamp=oversamp, grid_shape=grid_shape, kernel_width=kernel_width,
                    kernel_shape=kernel_shape, radius=PLAN_RADIUS)
    plan._home = home
    plan._dev = dev_fields
    plan._compiled = SimpleNamespace(
        device=dev, S=S, npts=npts, kw=kw, nper=nper, G=G, ndim=ndim, spatial=spatial,
        cell_ptr=cell_ptr, rep_src=rep_src, rep_s=rep_s, rep_ent=None, cell_unit=cell_unit, unit_cell=unit_cell,
        unit_rep=unit_rep, unit_hot=unit_hot, hot=hot, cell_vrow=cell_vrow, density=density, n_invalid=n_invalid,
        n_inv=n_inv,
        n_valid=[nper - v for v in n_inv], target_flat=target_idx, distances=distances, inv_perm_flat=inv_perm,
        interp_idx=None, entry_groups=None,
    )
    if s_shape:
        plan.stack_shape = s_shape
    return plan


def _attach_fft_plan(plan, image_shape=None):
    """FFT-plan scalars (reference _attach_fft_plan, _grog.py:918-998); the
    sorted arrays were produced by the plan kernels."""
    grid_shape = tuple(plan.grid_shape)
    if image_shape is None:
        image_shape = tuple(plan.shape)
    elif np.isscalar(image_shape):
        image_shape = len(grid_shape) * (int(image_shape),)
    plan.image_shape = tuple(int(s) for s in image_shape)
    plan.grid_size = int(np.prod(grid_shape))
    comp = plan._compiled
    plan.natural_shape = tuple(comp.spatial) + (comp.kw,)
    plan.n_samples = int(comp.nper)


# ---------------------------------------------------------------------------
# GrogInterpolator
# ---------------------------------------------------------------------------
class GrogInterpolator:
    """Data-driven GROG interpolator (reference ``GrogInterpolator``).

    Parameters are the reference's (_grog.py:302-312) plus ``device``, the
    HIP device that holds the plan and runs the kernels (default: current).
    """

    _interpolator_set = False
    _dataset_shape_set = False
    _data: list | None = None

    def __init__(self, shape, coords, oversamp=None, kernel_width: int = 2, kernel_shape: str = "circle",
                 time_map=None, image_shape=None, stack_shape: tuple = (), *, device=None):
        self._init_plan(_plan_finish(self._launch_plan(shape, coords, oversamp, kernel_width, kernel_shape, time_map,
                                                       stack_shape, device)), image_shape)

    @staticmethod
    def _launch_plan(shape, coords, oversamp=None, kernel_width: int = 2, kernel_shape: str = "circle",
                     time_map=None, stack_shape: tuple = (), device=None, defer: bool = False) -> _PendingPlan:
        coords_arr = coords if isinstance(coords, torch.Tensor) else np.asarray(coords)
        stack_eff = tuple(int(s) for s in (stack_shape or ()))
        if not stack_eff and coords_arr.ndim == 4:
            stack_eff = (int(coords_arr.shape[0]),)
        return _plan_launch(shape, coords_arr, oversamp, kernel_width, kernel_shape, time_map,
                            stack_shape=stack_eff, device=device, defer=defer)

    @classmethod
    def _from_pending(cls, pending: _PendingPlan, image_shape=None) -> "GrogInterpolator":
        """The interpolator of a plan launched with ``_launch_plan(...,
        defer=True)`` (waits for the plan compiler's summary)."""
        obj = cls.__new__(cls)
        obj._init_plan(_plan_finish(pending), image_shape)
        return obj

    def _init_plan(self, plan, image_shape) -> None:
        self.plan = plan
        self.device = self.plan._compiled.device
        _attach_fft_plan(self.plan, image_shape)
        self._programs = {}
        self._frame_off = {}

    # -- metadata / convenience (reference :337-376) -------------------------
    def metadata(self):
        oshape = np.ceil(np.asarray(self.plan.shape) * np.asarray(self.plan.oversamp)).astype(int).tolist()
        return SimpleNamespace(shape=tuple(self.plan.shape), oshape=tuple(oshape), target_idx=self.plan.target_idx,
                               weights=self.plan.weights, time_map=self.plan.time_map)

    def fft_plan(self, image_shape=None):
        if image_shape is not None and tuple(np.atleast_1d(image_shape)) != tuple(self.plan.image_shape):
            plan = copy.copy(self.plan)
            _attach_fft_plan(plan, image_shape)
            return plan
        return self.plan

    # -- table (reference :382-422) ------------------------------------------
    def calc_interp_table(self, train_data, lamda: float = 0.01, precision: int = 1, *, quant: int | None = None,
                          axis_order: str = "reference", keep_fp64: bool = False):
        """Build the operator table (FP64 on the GPU) and quantise every
        replica's shift to a table entry (bit-exact, in the plan kernels).

        ``quant=q`` quantises shifts to 1/q grid steps (table of (q+1)^ndim
        entries) instead of ``precision`` decimal digits.
        """
        comp = self.plan._compiled
        if comp.distances is None:
            raise RuntimeError("calc_interp_table: shifts were already released by a previous call")
        res = build_table(train_data, self.plan.radius, precision, lamda, quant=quant, device=comp.device,
                          axis_order=axis_order, keep_fp64=keep_fp64)
        self._attach_table(res, precision, quant)

    def _attach_table(self, res, precision: int = 1, quant: int | None = None, packed: torch.Tensor | None = None,
                      pack: bool = True):
        """Quantise every replica's shift against the fitted table ``res``
        and keep the table (``packed``: an already packed copy to share;
        ``pack=False``: pack the plain kernel's copy only when a plain launch
        first needs it -- the grouped modes never read it)."""
        comp = self.plan._compiled
        if comp.distances is None:
            raise RuntimeError("calc_interp_table: shifts were already released by a previous call")
        dev = comp.device
        pfac, stepsize, nsteps = quant_steps(self.plan.radius, precision, quant)
        if res.table.shape[0] != nsteps ** comp.ndim:
            raise RuntimeError("table size does not match the quantisation")
        K = int(res.table.shape[0])
        interp_idx = torch.empty(comp.S, comp.nper, dtype=torch.int64, device=dev)
        rep_ent = torch.zeros(comp.S, max(comp.nper, 1), dtype=torch.int32, device=dev)
        _lib.call("pg_plan_quantize", _lib.ptr(comp.distances), _lib.ptr(comp.inv_perm_flat), _lib.ptr(comp.n_invalid),
                  _lib.ptr(comp.target_flat), comp.S, comp.nper, comp.ndim, float(np.float32(pfac)),
                  float(np.float32(stepsize)), float(np.float32(self.plan.radius)), int(nsteps), K,
                  _lib.ptr(interp_idx), _lib.ptr(rep_ent), max(comp.nper, 1), _lib.stream(dev))
        comp.rep_ent = rep_ent
        comp.interp_idx = interp_idx
        comp.entry_groups = None
        comp.stack_groups = {}
        C = int(res.table.shape[1])
        comp.C = C
        comp.table = res.table
        comp.table64 = res.table64
        comp.powers64 = getattr(res, "powers64", None)
        comp.table_nsteps = int(nsteps)
        comp.table_axis_order = getattr(res, "axis_order", "reference")
        comp.table_host_steps = res.host_steps
        comp.table_packed = packed if C <= 64 else None
        if C <= 64 and packed is None and pack:
            self._pack_table()
        comp.table_bf16 = {}
        # release the shifts, as the reference does
        comp.distances = None
        self.plan._dev["distances"] = None
        self.plan.distances = None
        self._interp_kernel = res.table
        full = self.plan._dev["target_idx"].shape
        self._interp_idx = interp_idx.reshape(full)
        self._interpolator_set = True
        self._programs = {}

    def _pack_table(self) -> None:
        """The plain kernel's copy of the table (coil pairs, zero padded to
        the coil group)."""
        comp = self.plan._compiled
        K, C, dev = int(comp.table.shape[0]), comp.C, comp.device
        packed = torch.empty(int(_lib.load().pg_table_packed_elems(K, C)) * 4, dtype=torch.float32, device=dev)
        _lib.call("pg_table_pack_c64", _lib.ptr(comp.table), K, C, _lib.ptr(packed), _lib.stream(dev))
        comp.table_packed = packed

    def set_dataset_shape(self, shape):
        if self._dataset_shape_set:
            return
        self.dataset_shape = tuple(shape)
        self._dataset_shape_set = True

    # -- data layout helpers --------------------------------------------------
    def _split(self, data: torch.Tensor):
        """(*B, *S, C, *spatial) -> (batch_shape, C)."""
        comp = self.plan._compiled
        s_shape = tuple(self.plan.stack_shape)
        spatial = tuple(comp.spatial)
        ncoils_pos = data.ndim - len(spatial) - 1
        if ncoils_pos < len(s_shape):
            raise ValueError(f"data.ndim={data.ndim} too small for stack_shape={s_shape} + spatial={spatial} (+coils)")
        if s_shape:
            pref = tuple(int(s) for s in data.shape[ncoils_pos - len(s_shape):ncoils_pos])
            if pref != s_shape:
                raise ValueError(f"data stack prefix {pref} != plan.stack_shape {s_shape}")
        if tuple(int(s) for s in data.shape[ncoils_pos + 1:]) != spatial:
            raise ValueError(f"data spatial dims {tuple(data.shape[ncoils_pos + 1:])} != plan {spatial}")
        batch_shape = tuple(int(s) for s in data.shape[: ncoils_pos - len(s_shape)])
        return batch_shape, int(data.shape[ncoils_pos])

    def _to_device_c64(self, data: torch.Tensor, B: int, S: int, C: int) -> torch.Tensor:
        comp = self.plan._compiled
        return data.to(comp.device).to(torch.complex64).reshape(B, S, C, comp.npts).contiguous()

    def _frames(self, y: torch.Tensor, out: torch.Tensor, B: int, vbuf: torch.Tensor | None = None,
                v_row: int | None = None):
        """Frames of a (B, S) launch; ``y`` holds the per-frame sources
        (sample-major (N, C) rows for the fp32 mode) back to back."""
        comp = self.plan._compiled
        S, C, N, G = comp.S, comp.C, comp.npts, comp.G
        # bytes of one frame's sources: complex64 rows, or the 24-bit rows
        # of the BF16 modes (which only the reduction's frame table carries)
        y_frame = (y.numel() * y.element_size()) // max(B * S, 1)
        if vbuf is None and comp.table_packed is None and C <= 64:
            self._pack_table()
        frames = []
        for b in range(B):
            for s in range(S):
                f = _lib.Frame()
                f.y = y.data_ptr() + (b * S + s) * y_frame
                f.grid = out.data_ptr() + ((b * S + s) * C * G) * 8
                f.table = _lib.ptr(comp.table_packed) if comp.table_packed is not None else None
                stride = max(comp.nper, 1) * 4
                f.rep_src = comp.rep_src.data_ptr() + s * stride
                f.rep_ent = comp.rep_ent.data_ptr() + s * stride
                f.rep_s = comp.rep_s.data_ptr() + s * stride
                f.unit_cell = comp.unit_cell.data_ptr()
                f.unit_rep = comp.unit_rep.data_ptr() + s * 4
                f.cell_unit = comp.cell_unit.data_ptr() + s * G * 4
                if vbuf is not None:
                    # a frame's V: rmax rows of v_row bytes (complex64 or
                    # block-floating-point rows), frames back to back
                    off = (b * S + s) * self._rmax() * (v_row if v_row is not None else 8 * C)
                    f.vbuf = vbuf.data_ptr() + off
                hot = comp.hot[s]
                f.unit_hot = comp.unit_hot.data_ptr()
                f.hot_units = hot["hot_units"].data_ptr()
                f.hot_ptr = hot["hot_ptr"].data_ptr()
                f.n_hot = hot["n_hot"]
                f.n_hot_units = hot["n_hot_units"]
                f.cell_row = comp.cell_ptr[s].data_ptr()
                f.cell_vrow = comp.cell_vrow[s].data_ptr()
                f.n_src = N
                f.n_cell = G
                frames.append(f)
        return frames

    def _rmax(self) -> int:
        """Largest valid-replica count over the stack elements (rows of the
        grouped modes' intermediate buffer per frame)."""
        return max(1, max(self.plan._compiled.n_valid))

    def grid_program(self, y: torch.Tensor, out: torch.Tensor, B: int = 1, vbuf=None, n_blocks=None,
                     mode: str | None = None) -> GridProgram:
        comp = self.plan._compiled
        tc = mode_default_tile(mode, comp.C)
        tuned = getattr(self, "_tuned", {}).get(mode)
        if tuned is not None and n_blocks is None:
            tc, n_blocks = tuned
        v_row = sample_row_bytes(comp.C, v_format(mode)) if mode is not None and mode != "fp32" else None
        return GridProgram(self._frames(y, out, B, vbuf, v_row), comp.C, comp.G, comp.device, tc, n_blocks,
                           tile_opts=mode_tile_options(mode, comp.C))

    # -- grouped (BF16 MFMA) structures ---------------------------------------
    GROUP_CHUNK = 128  # replicas per MFMA work item (8 tiles of 16)

    def _entry_groups(self, defer: bool = False):
        """Per stack element: replicas of each table entry (stable radix sort of
        the entries) as (ent_ptr (K+1), ent_rep (R), items (n, 2) work list of
        (entry, first replica offset) chunks of GROUP_CHUNK).  ``defer``: built
        without a host read (the work list is launched over its capacity)."""
        comp = self.plan._compiled
        if comp.entry_groups is not None:
            return comp.entry_groups
        dev = comp.device
        K = int(comp.table.shape[0])
        groups = []
        for s in range(comp.S):
            R = comp.n_valid[s]
            groups.append(group_by_entry(comp.rep_ent[s, :max(R, 1)], R, K, dev, self.GROUP_CHUNK,
                                         comp.rep_src[s], comp.rep_s[s], defer=defer))
        comp.entry_groups = groups
        return groups

    def _stack_groups(self, B: int):
        """Entry groups of every (batch, stack element) frame of a device
        gridding call in one replica space: frame f = b * S + s owns replica
        rows [f * rmax, f * rmax + R_s) (the rows its V block has) and
        sample rows from f * N; the unused rows of a frame carry entry K,
        whose work items the apply skips.  None when the indices would not
        fit 32 bits."""
        comp = self.plan._compiled
        cache = getattr(comp, "stack_groups", None)
        if cache is None:
            cache = comp.stack_groups = {}
        if B in cache:
            return cache[B]
        dev, S, N = comp.device, comp.S, comp.npts
        K, rmax, F = int(comp.table.shape[0]), self._rmax(), B * S
        if F * max(N, rmax) >= 2 ** 31:
            cache[B] = None
            return None
        nv = torch.tensor(comp.n_valid, dtype=torch.int64, device=dev)
        live = torch.arange(rmax, device=dev)[None, :] < nv[:, None]
        ent = torch.where(live, comp.rep_ent[:, :rmax], K).to(torch.int32).repeat(B, 1)
        base = (torch.arange(F, dtype=torch.int64, device=dev) * N)[:, None]
        src = (comp.rep_src[:, :rmax].to(torch.int64).repeat(B, 1) + base).to(torch.int32)
        sw = comp.rep_s[:, :rmax].repeat(B, 1)
        grp = group_by_entry(ent.reshape(-1), F * rmax, K + 1, dev, self.GROUP_CHUNK, src.reshape(-1).contiguous(),
                             sw.reshape(-1).contiguous())
        cache[B] = grp
        return grp

    def _table_bf16(self, n_parts: int):
        comp = self.plan._compiled
        if n_parts not in comp.table_bf16:
            K, C = int(comp.table.shape[0]), comp.C
            n = int(_lib.load().pg_table_bf16_elems(K, C, n_parts))
            packed = torch.empty(n, dtype=torch.int16, device=comp.device)
            _lib.call("pg_table_pack_bf16", _lib.ptr(comp.table), K, C, n_parts, _lib.ptr(packed),
                      _lib.stream(comp.device))
            comp.table_bf16[n_parts] = packed
        return comp.table_bf16[n_parts]

    # -- gridding on device (no host transfer) --------------------------------
    def _transpose_all(self, y: torch.Tensor, yt: torch.Tensor, timer) -> None:
        """(B, S, C, N) coil-major -> per frame (N, C) sample-major rows
        (complex64, or the 24-bit format when ``yt`` is a byte buffer)."""
        comp = self.plan._compiled
        B, S, C, N = y.shape
        fmt = row_format_of(yt)
        st = _lib.stream(comp.device)
        F = B * S
        if F == 1:
            name = transpose_entry(fmt, False)
            e0 = timer.start() if timer is not None else None
            _lib.call(name, _lib.ptr(y), C, N, _lib.ptr(yt), st)
        else:
            # every (batch, stack element) block in one launch: equal-length
            # frames back to back, frame f's rows at sample offset f * N
            key = (F, N)
            off = getattr(self, "_frame_off", {}).get(key)
            if off is None:
                off = torch.arange(F + 1, dtype=torch.int64, device=comp.device) * N
                self._frame_off = {key: off}
            name = transpose_entry(fmt, True)
            e0 = timer.start() if timer is not None else None
            _lib.call(name, _lib.ptr(y), _lib.ptr(off), F, N, C, _lib.ptr(yt), st)
        if timer is not None:
            timer.stop(name, e0)

    def grid_device(self, y: torch.Tensor, mode: str = "fp32", out: torch.Tensor | None = None,
                    timer: KernelTimer | None = None, program_cache: dict | None = None) -> torch.Tensor:
        """Grid device-resident ``y`` (B, S, C, N) complex64 -> (B, S, C, G).

        Every mode first turns the coil-major sources into sample-major rows
        (one transpose pass), so a replica reads one contiguous row.
        mode "fp32": plain fused kernel (per-replica operator apply).
        "bf16" / "bf16x3": grouped complex GEMM on BF16 MFMA (1 or 3 split
        products, FP32 accumulate), then the same deterministic reduction;
        "bf16x3c": bf16x3 over block-floating-point sample rows (a 4-coil
        group's values within 2^-14 of its largest magnitude; half the
        bytes of complex64 rows).
        """
        if not self._interpolator_set:
            raise RuntimeError("GRAPPA kernels not set. Call calc_interp_table() first.")
        comp = self.plan._compiled
        if comp.C > 64:
            raise RuntimeError(f"fused gridding supports up to 64 coils, got {comp.C}")
        if mode not in GRID_MODES:
            raise ValueError(f"unknown gridding mode {mode!r}")
        B, S, C, N = y.shape
        if (S, C, N) != (comp.S, comp.C, comp.npts):
            raise ValueError(f"y shape {tuple(y.shape)} does not match plan (B, {comp.S}, {comp.C}, {comp.npts})")
        if not y.is_contiguous() or y.dtype != torch.complex64:
            raise ValueError("y must be a contiguous complex64 tensor")
        if out is None:
            out = torch.empty(B, S, C, comp.G, dtype=torch.complex64, device=comp.device)
        cache = program_cache if program_cache is not None else {}
        key = (mode, y.data_ptr(), out.data_ptr(), B)
        if key not in cache:
            yt = new_sample_rows(B * S * N, C, row_format(mode, C), comp.device)
            if mode == "fp32":
                cache[key] = (None, yt, self.grid_program(yt, out, B, mode=mode))
            else:
                vbuf = new_sample_rows(B * S * self._rmax(), C, v_format(mode), comp.device)
                cache[key] = (vbuf, yt, self.grid_program(yt, out, B, vbuf=vbuf, mode=mode))
        vbuf, yt, prog = cache[key]
        self._transpose_all(y, yt, timer)
        if mode == "fp32":
            prog.launch("pg_grid_fp32", timer)
            return out
        n_terms = mode_terms(mode)
        tbl = self._table_bf16(2 if n_terms == 3 else 1)
        groups = self._entry_groups()
        st = _lib.stream(comp.device)
        K = int(comp.table.shape[0])
        fmt = row_format_of(yt)
        rb = sample_row_bytes(C, fmt)
        name = apply_entry(fmt)
        sg = self._stack_groups(B) if B * S > 1 else None
        e0 = timer.start() if timer is not None else None
        if sg is not None:
            # every frame's replicas in one launch
            if sg.n_items > 0:
                _lib.call(name, yt.data_ptr(), _lib.ptr(tbl), _lib.ptr(sg.ent_ptr), _lib.ptr(sg.ent_rep),
                          _lib.ptr(sg.ent_src), _lib.ptr(sg.ent_s), _lib.ptr(sg.items), sg.n_items, K, C, n_terms,
                          vbuf.data_ptr(), st)
        else:
            for b in range(B):
                for s in range(S):
                    grp = groups[s]
                    if comp.n_valid[s] == 0 or grp.n_items == 0:
                        continue
                    yf = yt.data_ptr() + (b * S + s) * N * rb
                    vb = vbuf.data_ptr() + (b * S + s) * self._rmax() * sample_row_bytes(C, v_format(mode))
                    _lib.call(name, yf, _lib.ptr(tbl), _lib.ptr(grp.ent_ptr), _lib.ptr(grp.ent_rep),
                              _lib.ptr(grp.ent_src), _lib.ptr(grp.ent_s), _lib.ptr(grp.items), grp.n_items, K, C,
                              n_terms, vb, st)
        if timer is not None:
            timer.stop(name, e0)
        prog.launch(reduce_entry(mode), timer)
        return out

    def autotune(self, y: torch.Tensor, mode: str = "fp32", out: torch.Tensor | None = None,
                 program_cache: dict | None = None, reps: int = 5) -> dict:
        """Pick the gridding kernel's tile size and grid size for ``y`` on
        this device (timed on the real data; the bits do not change) and
        keep them for later ``grid_device`` calls with the same buffers and
        for new programs of this interpolator."""
        cache = program_cache if program_cache is not None else {}
        out = self.grid_device(y, mode, out, program_cache=cache)
        B = y.shape[0]
        _, _, prog = cache[(mode, y.data_ptr(), out.data_ptr(), B)]
        res = tune_program(prog, reduce_entry(mode), reps)
        self._tuned = getattr(self, "_tuned", {})
        self._tuned[mode] = (res["tile_cells"], res["n_blocks"])
        return res

    # -- forward interpolation (reference :437-555) ----------------------------
    def interpolate(self, data, shot_index=None, ret_image: bool = False, grid: bool = False, *, mode: str = "fp32"):
        if not self._interpolator_set:
            raise RuntimeError("GRAPPA kernels not set. Call calc_interp_table() first.")
        if grid and ret_image:
            raise ValueError("'grid' and 'ret_image' are mutually exclusive.")
        is_numpy = isinstance(data, np.ndarray)
        data_t = torch.as_tensor(data)
        src_device = data_t.device
        out_dtype = data_t.dtype if data_t.is_complex() else torch.complex64

        if shot_index is not None:
            if np.isscalar(shot_index):
                shot_index = (shot_index,)
            gridded = self._grid_shot(data_t, tuple(shot_index))
            if self._data is None:
                self._data = []
            self._data.append(gridded)
            return None

        plan = self.plan
        comp = plan._compiled
        batch_shape, C = self._split(data_t)
        if C != comp.C:
            raise ValueError(f"data has {C} coils but the table was trained for {comp.C}")
        B = int(np.prod(batch_shape)) if batch_shape else 1
        s_shape = tuple(plan.stack_shape)
        if data_t.dtype == torch.complex128:
            # complex128 data is computed in complex128, with the FP64 table
            if mode != "fp32":
                raise ValueError(f"mode {mode!r} computes in complex64; complex128 data is gridded in "
                                 "complex128 (mode='fp32')")
            out = self._interpolate_c128(data_t, batch_shape, B, C, ret_image, grid)
            if grid:
                kspace, mplan = out
                if is_numpy:
                    from ..operator._masked_fft import MaskedFFTPlan

                    return np.asarray(kspace.cpu()), MaskedFFTPlan(
                        grid_shape=mplan.grid_shape, image_shape=mplan.image_shape, stack_shape=s_shape,
                        mask=np.asarray(mplan.mask.cpu()), density=np.asarray(mplan.density.cpu()))
                mplan.mask = mplan.mask.to(src_device)
                mplan.density = mplan.density.to(src_device)
                return kspace.to(src_device), mplan
            out = out.to(src_device)
            return out.numpy() if is_numpy else out
        y = self._to_device_c64(data_t, B, comp.S, C)

        if ret_image or grid:
            g = self.grid_device(y, mode=mode)  # (B, S, C, G)
            gshape = tuple(plan.grid_shape)
            if ret_image:
                from .._base._fftc import ifft
                from .._utils import resize

                img = ifft(g.reshape(B, comp.S, C, *gshape), axes=tuple(range(-len(gshape), 0)))
                img = resize(img, (B, comp.S, C, *plan.image_shape))
                image = img.abs().square().sum(2).sqrt()
                image = image.reshape(*batch_shape, *s_shape, *plan.image_shape).to(src_device)
                return image.numpy() if is_numpy else image
            kspace = g.reshape(*batch_shape, *s_shape, C, *gshape).to(out_dtype).to(src_device)
            from ..operator._masked_fft import MaskedFFTPlan

            density = comp.density.reshape(*s_shape, *gshape)
            mplan = MaskedFFTPlan(grid_shape=gshape, image_shape=tuple(plan.image_shape), stack_shape=s_shape,
                                  mask=(density > 0).float().to(src_device), density=density.to(src_device))
            if is_numpy:
                numpy_plan = MaskedFFTPlan(grid_shape=gshape, image_shape=tuple(plan.image_shape), stack_shape=s_shape,
                                           mask=np.asarray(mplan.mask.cpu()), density=np.asarray(mplan.density.cpu()))
                return np.asarray(kspace.cpu()), numpy_plan
            return kspace, mplan

        sparse = torch.empty(B, comp.S, C, comp.nper, dtype=torch.complex64, device=comp.device)
        K = int(comp.table.shape[0])
        # every (batch, stack element) block in one launch
        _lib.call("pg_sparse_apply_stack_c64", _lib.ptr(y.contiguous()), _lib.ptr(comp.interp_idx.contiguous()),
                  _lib.ptr(comp.table), _lib.ptr(sparse), B, comp.S, C, comp.npts, comp.kw, K,
                  _lib.stream(comp.device))
        out = sparse.reshape(*batch_shape, *s_shape, C, comp.nper).to(out_dtype).to(src_device)
        return out.numpy() if is_numpy else out

    # -- complex128 (the reference keeps the data's precision end to end,
    # csrc/torch/grog_interp.cpp:112-149, calib/_grappa.py:62-72) ----------------
    def _table64(self) -> torch.Tensor:
        """The (K, C, C) complex128 table: kept at fit time or assembled
        (once) from the FP64 per-axis powers of the fit."""
        from ._grappa import compose_table

        comp = self.plan._compiled
        if getattr(comp, "table64", None) is None:
            if getattr(comp, "powers64", None) is None:
                raise RuntimeError("complex128 data needs the FP64 table: this table was set without its FP64 "
                                   "powers (fit it with calc_interp_table)")
            P = torch.stack([g for g in comp.powers64 if g is not None]).to(comp.device)
            comp.table64 = compose_table(P, comp.ndim, comp.table_nsteps, comp.table_axis_order, keep_fp64=True)[1]
        return comp.table64

    def _apply_c128(self, y: torch.Tensor, s: int) -> torch.Tensor:
        """Every candidate replica of stack element s, in complex128, the
        reference flow (_grog.py:668-695): y (B, C, N) -> (B, C, N * kw)."""
        comp = self.plan._compiled
        B, C, N = y.shape
        kw = comp.kw
        rows = y.permute(2, 0, 1).repeat_interleave(kw, dim=0).contiguous()  # (N * kw, B, C)
        T = self._table64()
        out = torch.empty_like(rows)
        idx = comp.interp_idx[s].contiguous()
        _lib.call("pg_grog_interpolate_c128", _lib.ptr(rows), _lib.ptr(idx), _lib.ptr(T), _lib.ptr(out), N * kw, B,
                  C, int(T.shape[0]), _lib.stream(comp.device))
        return out.permute(1, 2, 0)

    def _interpolate_c128(self, data: torch.Tensor, batch_shape, B: int, C: int, ret_image: bool, grid: bool):
        plan = self.plan
        comp = plan._compiled
        dev = comp.device
        S, G = comp.S, comp.G
        s_shape = tuple(plan.stack_shape)
        y = data.to(dev).reshape(B, S, C, comp.npts)
        sparse = torch.empty(B, S, C, comp.nper, dtype=torch.complex128, device=dev)
        for s in range(S):
            sparse[:, s] = self._apply_c128(y[:, s].contiguous(), s)
        if not (grid or ret_image):
            return sparse.reshape(*batch_shape, *s_shape, C, comp.nper)
        # pre-weights, sorted order, then the order-fixed scatter of every
        # (batch, coil) row with the sqrt weights (_grog.py:763-821)
        idx = self.plan._dev["indices"].reshape(S, comp.nper)
        sqw = self.plan._dev["sqrt_weights"].reshape(S, comp.nper).double()
        sp = self.plan._dev["sort_perm"].reshape(S, comp.nper)
        ip = self.plan._dev["inv_perm"].reshape(S, comp.nper)
        g = torch.zeros(B, S, C, G, dtype=torch.complex128, device=dev)
        st = _lib.stream(dev)
        for s in range(S):
            pre_w = sqw[s][ip[s]].to(torch.complex128)
            rows = (sparse[:, s] * pre_w)[:, :, sp[s]].reshape(B * C, comp.nper).contiguous()
            gs = torch.zeros(B * C, G, dtype=torch.complex128, device=dev)
            w = sqw[s].contiguous()
            ix = idx[s].contiguous()
            for r0 in range(0, B * C, 65535):
                r1 = min(r0 + 65535, B * C)
                _lib.call("pg_scatter_add_rows_c128", _lib.ptr(gs[r0:r1]), G, _lib.ptr(rows[r0:r1]), _lib.ptr(ix),
                          _lib.ptr(w), comp.nper, r1 - r0, st)
            g[:, s] = gs.reshape(B, C, G)
        gshape = tuple(plan.grid_shape)
        if ret_image:
            from .._base._fftc import ifft
            from .._utils import resize

            img = ifft(g.reshape(B, S, C, *gshape), axes=tuple(range(-len(gshape), 0)))
            img = resize(img, (B, S, C, *plan.image_shape))
            return img.abs().square().sum(2).sqrt().reshape(*batch_shape, *s_shape, *plan.image_shape)
        from ..operator._masked_fft import MaskedFFTPlan

        density = comp.density.reshape(*s_shape, *gshape)
        mplan = MaskedFFTPlan(grid_shape=gshape, image_shape=tuple(plan.image_shape), stack_shape=s_shape,
                              mask=(density > 0).float(), density=density)
        return g.reshape(*batch_shape, *s_shape, C, *gshape), mplan

    def __call__(self, data, shot_index=None, ret_image=False, grid=False):
        return self.interpolate(data, shot_index, ret_image, grid)

    # -- shot-by-shot (reference :697-731, 904-912) ---------------------------
    def _grid_shot(self, data: torch.Tensor, shot_index: tuple) -> torch.Tensor:
        comp = self.plan._compiled
        dev = comp.device
        tidx = self.plan._dev["target_idx"]
        w_all = self.plan._dev["weights"]
        iidx = self._interp_idx
        t_shot = tidx[shot_index].reshape(-1).contiguous()
        w_shot = w_all[shot_index].reshape(-1).contiguous()
        e_shot = iidx[shot_index].reshape(-1).contiguous()
        kw = comp.kw
        npts = t_shot.numel() // kw
        C = int(data.shape[0])
        if data.dtype == torch.complex128:
            # the reference shot flow in complex128 (operator, weight, serial
            # scatter in natural order within a cell)
            y = data.to(dev).reshape(C, npts).contiguous()
            rows = y.T.repeat_interleave(kw, dim=0).reshape(npts * kw, 1, C).contiguous()
            T = self._table64()
            applied = torch.empty_like(rows)
            _lib.call("pg_grog_interpolate_c128", _lib.ptr(rows), _lib.ptr(e_shot), _lib.ptr(T), _lib.ptr(applied),
                      npts * kw, 1, C, int(T.shape[0]), _lib.stream(dev))
            t_sorted, order = torch.sort(t_shot, stable=True)
            vals = applied[:, 0, :].T[:, order].contiguous()
            out = torch.zeros(C, comp.G, dtype=torch.complex128, device=dev)
            _lib.call("pg_scatter_add_rows_c128", _lib.ptr(out), comp.G, _lib.ptr(vals), _lib.ptr(t_sorted),
                      _lib.ptr(w_shot[order].double().contiguous()), npts * kw, C, _lib.stream(dev))
            return out.to(data.device)
        y = data.to(dev).to(torch.complex64).reshape(1, C, npts).contiguous()
        applied = torch.empty(1, C, npts * kw, dtype=torch.complex64, device=dev)
        st = _lib.stream(dev)
        _lib.call("pg_sparse_apply_c64", _lib.ptr(y), _lib.ptr(e_shot), _lib.ptr(comp.table), _lib.ptr(applied), 1, C,
                  npts, kw, int(comp.table.shape[0]), st)
        out = torch.zeros(C, comp.G, dtype=torch.complex64, device=dev)
        # sorted by cell, replicas of a cell in natural order (a stable sort):
        # the scatter then folds each cell in the order of the reference's
        # serial scatter_add_, the same bits on every call
        t_sorted, order = torch.sort(t_shot, stable=True)
        applied = applied[:, :, order].contiguous()
        w_sorted = w_shot[order].contiguous()
        # every coil in one launch
        _lib.call("pg_scatter_add_rows_c64", _lib.ptr(out), comp.G, _lib.ptr(applied), _lib.ptr(t_sorted),
                  _lib.ptr(w_sorted), npts * kw, C, st)
        return out.to(data.device)

    def collect_shots(self) -> torch.Tensor:
        if self._data is None or len(self._data) == 0:
            raise RuntimeError("No shots accumulated.")
        output = torch.stack(self._data, dim=0).sum(dim=0)
        self._data = None
        return output.reshape(output.shape[0], *self.plan.grid_shape)

    # -- statistics used by the benchmark ---------------------------------------
    def work_stats(self) -> dict:
        comp = self.plan._compiled
        return {
            "S": comp.S, "npts": comp.npts, "kw": comp.kw, "G": comp.G, "C": getattr(comp, "C", None),
            "replicas_valid": int(sum(comp.n_valid)), "units": int(comp.cell_unit[-1].item()),
            "table_entries": int(comp.table.shape[0]) if getattr(comp, "table", None) is not None else None,
        }


def _free():
    gc.collect()
    if torch.cuda.is_available():
        torch.cuda.empty_cache()


_ = ctypes  # ctypes structures are built through _lib
