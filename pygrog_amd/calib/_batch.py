"""Many frames of uneven sample counts sharing one operator table, gridded
together on one GPU.

The reference grids a dynamic series frame by frame, one
``GrogInterpolator`` per frame (src/pygrog/calib/_grog.py:302-331 for the
plan, :382-422 for the table, :736-902 for ``interpolate(grid=True)``).
Here every frame is such an interpolator (same plan compiler, same
quantisation, the table fitted once and shared), and one call grids all of
them with a fixed, small number of launches whatever the frame count:

  1. ``pg_transpose_frames_c64`` (``_s24`` for the BF16 modes up to 32 coils:
     24-bit rows; ``_bfp`` for bf16x3c: block-floating-point rows) every
     frame's (C, N_f) block to (N_f, C) rows;
  2. fp32:    ``pg_grid_fp32`` over all frames' tiles (plain per-replica apply);
     bf16(x3, x3c): ``pg_apply_grouped_bf16`` once over the replicas of all frames
              grouped by table entry (the table is shared, so groups are F
              times larger than one frame's), then ``pg_grid_reduce_v`` over
              all frames' tiles.

``capture`` records that sequence once in a HIP graph (torch.cuda.CUDAGraph)
and replays it with one launch.  ``autotune`` times the launch shapes of the
gridding kernel on the real data and keeps the fastest.

Every frame's grid is bitwise the grid its own interpolator produces with
the same mode: the per-replica values do not depend on which other replicas
share a launch or an MFMA tile, and every cell folds its replicas in the
same canonical order.
"""

from __future__ import annotations

__all__ = ["GrogFrameBatch"]

import numpy as np
import torch

from .. import _lib
from ._grappa import build_table
from ._grog import (GRID_MODES, GridProgram, GrogInterpolator, KernelTimer, apply_entry, group_by_entry,
                    mode_default_tile, mode_terms, mode_tile_options, new_sample_rows, reduce_entry, row_format,
                    row_format_of, sample_row_bytes, transpose_entry, tune_program, v_format)

_TUNE_CACHE: dict = {}


class GrogFrameBatch:
    """``F`` frames on one grid, each with its own trajectory ``coords[f]``
    of shape (..., ndim) (any number of samples), one shared table."""

    def __init__(self, shape, coords, oversamp=None, kernel_width: int = 2, kernel_shape: str = "circle",
                 image_shape=None, *, device=None):
        if len(coords) == 0:
            raise ValueError("GrogFrameBatch needs at least one frame")
        if len(coords) >= 65536:
            raise ValueError("GrogFrameBatch: at most 65535 frames per batch")
        self.frames = [GrogInterpolator(shape, c, oversamp, kernel_width, kernel_shape, image_shape=image_shape,
                                        device=device) for c in coords]
        f0 = self.frames[0]
        self.device = f0.device
        self.grid_shape = tuple(int(s) for s in f0.plan.grid_shape)
        for f in self.frames[1:]:
            if tuple(int(s) for s in f.plan.grid_shape) != self.grid_shape:
                raise ValueError("all frames must share one grid")
        self.F = len(self.frames)
        self.G = int(f0.plan._compiled.G)
        self.n_samples = np.array([f.plan._compiled.npts for f in self.frames], dtype=np.int64)
        self.sample_off = np.concatenate([[0], np.cumsum(self.n_samples)]).astype(np.int64)
        self._off_dev = torch.as_tensor(self.sample_off, device=self.device)
        self.C = None
        self._merged = None
        self._programs = {}
        self.tuned = {}

    # -- table ---------------------------------------------------------------
    def calc_interp_table(self, train_data, lamda: float = 0.01, precision: int = 1, *, quant: int | None = None,
                          axis_order: str = "reference", keep_fp64: bool = False):
        """Fit the table once (FP64 on the GPU) and quantise every frame's
        shifts against it (bit-exact, per frame)."""
        f0 = self.frames[0]
        res = build_table(train_data, f0.plan.radius, precision, lamda, quant=quant, device=self.device,
                          axis_order=axis_order, keep_fp64=keep_fp64)
        f0._attach_table(res, precision, quant)
        packed = f0.plan._compiled.table_packed
        for f in self.frames[1:]:
            f._attach_table(res, precision, quant, packed=packed)
        self.C = int(res.table.shape[1])
        self.K = int(res.table.shape[0])
        self._merged = None
        self._programs = {}
        return res

    # -- data layout -----------------------------------------------------------
    def pack(self, data) -> torch.Tensor:
        """List of per-frame (C, N_f) k-space -> one flat complex64 device
        tensor, frame f at element offset C * sample_off[f]."""
        if len(data) != self.F:
            raise ValueError(f"expected {self.F} frames, got {len(data)}")
        parts = []
        for f, d in enumerate(data):
            t = torch.as_tensor(d)
            if t.numel() != self.C * int(self.n_samples[f]):
                raise ValueError(f"frame {f}: expected {self.C} x {int(self.n_samples[f])} samples, got {tuple(t.shape)}")
            parts.append(t.reshape(-1).to(self.device, torch.complex64))
        return torch.cat(parts)

    def replicas(self) -> np.ndarray:
        return np.array([f.plan._compiled.n_valid[0] for f in self.frames], dtype=np.int64)

    def _merged_groups(self):
        """All frames' valid replicas in one global index space (frame f's
        replicas at offset sum_{g<f} R_g), grouped by table entry."""
        if self._merged is not None:
            return self._merged
        R = self.replicas()
        rep_off = np.concatenate([[0], np.cumsum(R)]).astype(np.int64)
        src, sw, ent = [], [], []
        for f, fr in enumerate(self.frames):
            comp = fr.plan._compiled
            r = int(R[f])
            src.append(comp.rep_src[0, :r].to(torch.int64) + int(self.sample_off[f]))
            sw.append(comp.rep_s[0, :r])
            ent.append(comp.rep_ent[0, :r])
        Rt = int(rep_off[-1])
        if int(self.sample_off[-1]) >= 2 ** 31 or Rt >= 2 ** 31:
            raise ValueError("GrogFrameBatch: more than 2^31 samples or replicas in one batch")
        rep_src = torch.cat(src).to(torch.int32).contiguous()
        rep_s = torch.cat(sw).contiguous()
        rep_ent = torch.cat(ent).contiguous()
        grp = group_by_entry(rep_ent, Rt, self.K, self.device, GrogInterpolator.GROUP_CHUNK, rep_src, rep_s)
        self._merged = {"rep_off": rep_off, "R": Rt, "groups": grp, "n_items": grp.n_items}
        return self._merged

    def _program(self, mode: str, y: torch.Tensor, out: torch.Tensor):
        key = (mode, y.data_ptr(), out.data_ptr())
        if key in self._programs:
            return self._programs[key]
        C = self.C
        yt = new_sample_rows(int(self.sample_off[-1]), C, row_format(mode, C), self.device)
        rb = sample_row_bytes(C, row_format_of(yt))
        tc = mode_default_tile(mode, C)
        frames = []
        vbuf = None
        if mode != "fp32":
            m = self._merged_groups()
            vbuf = new_sample_rows(max(m["R"], 1), C, v_format(mode), self.device)
            vrow = sample_row_bytes(C, v_format(mode)) // vbuf.element_size()
        for f, fr in enumerate(self.frames):
            yv = yt[(rb // yt.element_size()) * int(self.sample_off[f]):]
            vv = None if vbuf is None else vbuf[vrow * int(self._merged["rep_off"][f]):]
            frames += fr._frames(yv, out[f], 1, vv)
        tune = self.tuned.get(mode, {})
        prog = GridProgram(frames, C, self.G, self.device, tune.get("tile_cells", tc), tune.get("n_blocks"),
                           tile_opts=mode_tile_options(mode, C))
        self._programs[key] = (yt, vbuf, prog)
        return self._programs[key]

    def _check(self, y: torch.Tensor, out: torch.Tensor, mode: str):
        if self.C is None:
            raise RuntimeError("GRAPPA kernels not set. Call calc_interp_table() first.")
        if mode not in GRID_MODES:
            raise ValueError(f"unknown gridding mode {mode!r}")
        if self.C > 64:
            raise RuntimeError(f"fused gridding supports up to 64 coils, got {self.C}")
        if y.device != self.device or y.dtype != torch.complex64 or not y.is_contiguous():
            raise ValueError("y must be a contiguous complex64 tensor on the batch's device")
        if y.numel() != self.C * int(self.sample_off[-1]):
            raise ValueError(f"y holds {y.numel()} values, expected {self.C} x {int(self.sample_off[-1])}")
        if out.shape[0] != self.F or out.numel() != self.F * self.C * self.G or not out.is_contiguous() \
                or out.dtype != torch.complex64 or out.device != self.device:
            raise ValueError(f"out must be a contiguous complex64 ({self.F}, {self.C}, *{self.grid_shape}) tensor")

    def _run(self, mode: str, y: torch.Tensor, out: torch.Tensor, timer: KernelTimer | None = None):
        yt, vbuf, prog = self._program(mode, y, out)
        st = _lib.stream(self.device)
        fmt = row_format_of(yt)
        name = transpose_entry(fmt, True)
        e0 = timer.start() if timer is not None else None
        _lib.call(name, _lib.ptr(y), _lib.ptr(self._off_dev), self.F, int(self.n_samples.max()),
                  self.C, _lib.ptr(yt), st)
        if timer is not None:
            timer.stop(name, e0)
        if mode == "fp32":
            prog.launch("pg_grid_fp32", timer)
            return
        m = self._merged_groups()
        n_terms = mode_terms(mode)
        tbl = self.frames[0]._table_bf16(2 if n_terms == 3 else 1)
        name = apply_entry(fmt)
        e0 = timer.start() if timer is not None else None
        if m["n_items"] > 0:
            grp = m["groups"]
            _lib.call(name, _lib.ptr(yt), _lib.ptr(tbl), _lib.ptr(grp.ent_ptr),
                      _lib.ptr(grp.ent_rep), _lib.ptr(grp.ent_src), _lib.ptr(grp.ent_s), _lib.ptr(grp.items),
                      grp.n_items, self.K, self.C, n_terms, _lib.ptr(vbuf), st)
        if timer is not None:
            timer.stop(name, e0)
        prog.launch(reduce_entry(mode), timer)

    def new_output(self) -> torch.Tensor:
        return torch.empty(self.F, self.C, *self.grid_shape, dtype=torch.complex64, device=self.device)

    # -- public gridding -------------------------------------------------------
    def grid(self, y, mode: str = "fp32", out: torch.Tensor | None = None,
             timer: KernelTimer | None = None) -> torch.Tensor:
        """Grid all frames: ``y`` packed (see ``pack``) or a list of per-frame
        (C, N_f) arrays -> (F, C, *grid_shape) complex64 on the device."""
        if not isinstance(y, torch.Tensor):
            y = self.pack(y)
        if out is None:
            out = self.new_output()
        self._check(y, out, mode)
        self._run(mode, y, out, timer)
        return out

    def capture(self, y: torch.Tensor, mode: str = "fp32", out: torch.Tensor | None = None):
        """Record the whole batch gridding once in a HIP graph; returns
        (graph, out): ``graph.replay()`` re-grids whatever ``y`` holds into
        ``out`` with one launch."""
        if out is None:
            out = self.new_output()
        self._check(y, out, mode)
        self._run(mode, y, out)  # builds every structure outside the capture
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self._run(mode, y, out)
        return graph, out

    def autotune(self, y: torch.Tensor, mode: str = "fp32", out: torch.Tensor | None = None, reps: int = 5,
                 candidates=None) -> dict:
        """Time the gridding kernel over its compiled tile sizes and
        candidate persistent grid sizes on this batch's data and keep the
        fastest (the shape never changes the result bits); results are
        cached per (mode, coils, cells, frames, replicas).  Capture after
        tuning."""
        if out is None:
            out = self.new_output()
        self._check(y, out, mode)
        key = (mode, self.C, self.G, self.F, int(self.replicas().sum()))
        yt, vbuf, prog = self._program(mode, y, out)
        if key not in _TUNE_CACHE:
            self._run(mode, y, out)
            _TUNE_CACHE[key] = tune_program(prog, reduce_entry(mode), reps, candidates)
        self.tuned[mode] = dict(_TUNE_CACHE[key])
        for k, (yt_, vbuf_, pr) in self._programs.items():
            if k[0] == mode:
                pr.set_shape(self.tuned[mode]["tile_cells"], self.tuned[mode]["n_blocks"])
        return self.tuned[mode]

    # -- per-frame views -------------------------------------------------------
    def density(self, f: int) -> torch.Tensor:
        comp = self.frames[f].plan._compiled
        return comp.density.reshape(self.grid_shape)

    def work_stats(self) -> dict:
        R = self.replicas()
        return {"F": self.F, "C": self.C, "G": self.G, "samples": int(self.sample_off[-1]), "replicas_valid": int(R.sum()),
                "units": int(sum(int(f.plan._compiled.cell_unit[-1].item()) for f in self.frames)),
                "table_entries": self.K}
