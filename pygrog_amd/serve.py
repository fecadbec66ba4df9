"""Serving loop for GROG gridding jobs on one GPU: HBM-budget admission,
per-job preparation overlapped on several streams, and continuous batching
of the gridding launches.

A request is one reconstruction's gridding: a trajectory, its own
calibration region (ACR) and its k-space, as a user of the reference would
pass to ``GrogInterpolator(shape, coords, ...)``,
``calc_interp_table(acr, ...)`` and ``interpolate(data, grid=True)``
(src/pygrog/calib/_grog.py:302-331, 382-422, 437-555).  The loop:

  * admits arrived requests in arrival order while the HBM budget (an
    up-front estimate of every buffer the job will hold) and a preparation
    slot are free;
  * prepares each admitted job -- GPU plan compiler and FP64 table fit -- on
    its slot's own stream, from a worker thread (the fit is solved on the
    host while the plan kernels run), so preparations of different jobs
    overlap each other and the gridding on the device;
  * every time the gridding stream has room, gathers all prepared jobs
    (continuous batching: jobs that became ready while earlier batches were
    on the GPU join the next one) and grids them together: one transpose per
    job, then one fused launch over all frames of the same coil count and
    grid size (each frame carries its own table), or for the BF16 modes one
    grouped MFMA apply per job and one shared reduction launch;
  * retires a batch when its event completes, records each job's latency
    (arrival to gridded k-space resident on the device) and releases its
    budget.

A job's result is bitwise the one its own interpolator produces alone with
the same mode: every cell folds its replicas in a fixed order whatever
launch the job shares.

Failures stay with their request.  A malformed job (k-space that does not
match its trajectory or calibration, a stacked trajectory, too many coils)
and a job whose estimate exceeds the whole HBM budget are rejected before
admission; a job whose preparation raises, or whose buffers cannot be
allocated when its batch is launched, is marked failed and its budget and
slot are released.  Every other job is served as if it had not been there.
Each job records ``status`` ("done", "rejected" or "failed") and ``error``.
"""

from __future__ import annotations

__all__ = ["GridJob", "GrogServer", "estimate_job_bytes", "poisson_arrivals", "serve_loop", "validate_job"]

import concurrent.futures as cf
import dataclasses
import gc
import time

import numpy as np
import torch

from . import _lib
from .calib._grappa import build_table_finish, build_table_launch
from .calib._grog import (GRID_MODES, PLAN_RADIUS, GridProgram, GrogInterpolator, _default_oversamp,
                          _kernel_offsets, apply_entry, mode_default_tile, mode_terms, mode_tile_options,
                          reduce_entry, row_format, sample_row_bytes, transpose_entry, v_format)


@dataclasses.dataclass(eq=False)
class GridJob:
    """One gridding request (compared and hashed by identity).  ``kspace``
    (C, N) complex64 and ``calib`` (C, *acr) are device tensors (the request
    payload, staged on the GPU); ``coords`` is the trajectory (..., ndim),
    float32, a host array or a device tensor staged with the payload."""

    shape: tuple
    coords: np.ndarray
    calib: torch.Tensor
    kspace: torch.Tensor
    kernel_width: int = 2
    kernel_shape: str = "circle"
    oversamp: object = None
    lamda: float = 1e-3
    precision: int = 1
    quant: int | None = None
    arrival: float = 0.0
    job_id: int = 0
    # filled by the server
    t_admit: float | None = None
    t_ready: float | None = None
    t_done: float | None = None
    est_bytes: int = 0
    status: str | None = None
    error: str | None = None
    result: torch.Tensor | None = None
    interp: GrogInterpolator | None = None
    _ready_event: torch.cuda.Event | None = None

    @property
    def coils(self) -> int:
        return int(self.kspace.shape[0])

    @property
    def samples(self) -> int:
        return int(self.kspace.shape[-1])

    @property
    def latency(self) -> float | None:
        return None if self.t_done is None else self.t_done - self.arrival


def estimate_job_bytes(job: GridJob, mode: str) -> int:
    """Upper estimate of the device memory a job holds from admission to
    completion: plan (per candidate replica and per cell), table and its
    packed copies, the sample-major copy of the k-space, the grid and, for
    the grouped modes, the applied replicas."""
    ndim = len(job.shape)
    os_ = _default_oversamp(ndim, job.oversamp)
    grid = [int(np.ceil(s * o)) for s, o in zip(job.shape, os_)]
    G = int(np.prod(grid))
    kw = int(len(_kernel_offsets(ndim, job.kernel_width, job.kernel_shape)))
    N, C = job.samples, job.coils
    if job.quant is not None:
        nsteps = job.quant + 1
    else:
        nsteps = 10 ** job.precision + 1
    K = nsteps ** ndim
    # plan fields + records + sort buffers + the unit arrays' upper bound
    per_rep = 8 + 4 * ndim + 4 + 8 + 4 + 8 + 8 + 4 * 3 + 8 + 16 + 16
    plan = N * kw * per_rep + G * (4 * 4 + 4 + 4)
    table = K * C * C * (8 + 8 + 8 + 2 * 8 * 2)  # c64, fp64 pair, fp32 packed, bf16 hi/lo real form
    data = 2 * 8 * C * N + 8 * C * G
    vbuf = 8 * C * N * (ndim if mode != "fp32" else 0)
    return int(1.25 * (plan + table + data + vbuf)) + (16 << 20)


def poisson_arrivals(n: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times (s) of a Poisson process of ``rate`` jobs/s."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def validate_job(job: GridJob) -> str | None:
    """Why ``job`` cannot be served, or None.  Checks what the gridding
    launch assumes: one trajectory (no stack), k-space (C, N) complex with
    N the trajectory's sample count and C the calibration's coil count,
    at most 64 coils, a known kernel shape."""
    try:
        shape = tuple(int(s) for s in job.shape)
    except (TypeError, ValueError):
        return f"grid shape {job.shape!r} is not a tuple of integers"
    ndim = len(shape)
    if ndim not in (2, 3) or min(shape) < 1:
        return f"grid shape {shape} must have 2 or 3 positive axes"
    cshape = tuple(int(v) for v in (job.coords.shape if hasattr(job.coords, "shape") else np.shape(job.coords)))
    if len(cshape) < 2 or cshape[-1] != ndim:
        return f"coords shape {cshape} does not end in the grid's {ndim} axes"
    if len(cshape) > 3:
        return (f"coords shape {cshape} is a stack of trajectories; a serving job grids one "
                "trajectory (submit one job per stack element)")
    n = int(np.prod(cshape[:-1]))
    y = job.kspace
    if not isinstance(y, torch.Tensor) or y.ndim != 2 or not y.is_complex():
        return "kspace must be a complex tensor of shape (coils, samples)"
    C = int(y.shape[0])
    if int(y.shape[1]) != n:
        return f"kspace has {int(y.shape[1])} samples per coil but the trajectory has {n}"
    if not isinstance(job.calib, torch.Tensor) or job.calib.ndim != ndim + 1:
        return f"calib must be a tensor (coils, *acr) with {ndim} spatial axes"
    if int(job.calib.shape[0]) != C:
        return f"kspace has {C} coils but the calibration has {int(job.calib.shape[0])}"
    if not 1 <= C <= 64:
        return f"fused gridding supports 1 to 64 coils, got {C}"
    try:
        kw = len(_kernel_offsets(ndim, int(job.kernel_width), job.kernel_shape))
    except (ValueError, TypeError) as exc:
        return f"kernel: {exc}"
    if kw == 0:
        return f"kernel: width {job.kernel_width} gives no neighbours"
    return None


def serve_loop(jobs: list, *, budget: int, slots: int, max_inflight: int, estimate, prepare, launch,
               submit, timeout: float = 600.0, on_retire=None) -> dict:
    """The scheduling loop of a serving run, with the device work passed in.

    ``estimate(job)`` -> bytes; ``prepare(job, slot)`` runs through
    ``submit(fn, *args)`` (an executor's submit) and may raise: the job
    fails alone.  ``launch(batch)`` -> ``(handle, keep, failed)`` starts one
    gridding batch (``handle.query()`` is true once it has completed) and
    returns the jobs it could not launch, ``{job: message}``; ``keep`` is
    held until completion.  ``on_retire(job)`` runs for each completed job.
    Jobs that fail validation or exceed the whole budget are rejected
    before admission.  Returns the bookkeeping lists (done, rejected,
    failed, batch sizes, peak admitted bytes, makespan); each job's
    ``status`` and ``error`` are set."""
    t0 = time.perf_counter()
    pending, rejected = [], []
    for j in sorted(jobs, key=lambda j: j.arrival):
        j.status, j.error = None, None
        why = validate_job(j)
        if why is None:
            j.est_bytes = int(estimate(j))
            if j.est_bytes > budget:
                why = (f"job {j.job_id} needs ~{j.est_bytes / 1e9:.2f} GB, above the budget "
                       f"{budget / 1e9:.2f} GB")
        if why is not None:
            j.status, j.error = "rejected", why
            rejected.append(j)
        else:
            pending.append(j)
    n_served = len(pending)
    used = 0
    peak_used = 0
    preparing = {}
    ready, inflight, done, failed, batches = [], [], [], [], []
    free_slots = list(range(slots))

    def fail(job, msg):
        nonlocal used
        job.status, job.error = "failed", msg
        job.t_done = None
        job.result = None
        job.interp = None
        used -= job.est_bytes
        failed.append(job)

    try:
        while len(done) + len(failed) < n_served:
            now = time.perf_counter() - t0
            if now > timeout:
                raise RuntimeError(f"serving run exceeded {timeout:.0f} s")
            progressed = False
            # admit in arrival order while budget and a slot are free
            while pending and pending[0].arrival <= now and free_slots and used + pending[0].est_bytes <= budget:
                job = pending.pop(0)
                slot = free_slots.pop(0)
                used += job.est_bytes
                peak_used = max(peak_used, used)
                job.t_admit = now
                preparing[submit(prepare, job, slot)] = (job, slot)
                progressed = True
            # collect finished preparations; a preparation error fails its job only
            for fut in [f for f in preparing if f.done()]:
                job, slot = preparing.pop(fut)
                free_slots.append(slot)
                exc = fut.exception()
                if exc is not None:
                    fail(job, f"preparation failed: {type(exc).__name__}: {exc}")
                else:
                    job.t_ready = time.perf_counter() - t0
                    ready.append(job)
                progressed = True
            # continuous batching: everything ready joins the next launch
            if ready and len(inflight) < max_inflight:
                batch, ready = ready, []
                handle, keep, bad = launch(batch)
                for job, msg in bad.items():
                    fail(job, msg)
                batch = [j for j in batch if j not in bad]
                if batch:
                    inflight.append((handle, batch, keep))
                    batches.append(len(batch))
                progressed = True
            # retire completed batches
            still = []
            for handle, batch, keep in inflight:
                if handle.query():
                    t = time.perf_counter() - t0
                    for job in batch:
                        job.t_done = t
                        job.status = "done"
                        used -= job.est_bytes
                        if on_retire is not None:
                            on_retire(job)
                        done.append(job)
                    progressed = True
                else:
                    still.append((handle, batch, keep))
            inflight = still
            if not progressed:
                # wake as soon as a preparation ends (or after 0.2 ms, for
                # arrivals and batch completions)
                if preparing:
                    cf.wait(list(preparing), timeout=0.0002, return_when=cf.FIRST_COMPLETED)
                else:
                    time.sleep(0.0002)
    finally:
        # a failed run leaves no preparation behind
        cf.wait(list(preparing))
    makespan = max(j.t_done for j in done) if done else 0.0
    return {"done": done, "rejected": rejected, "failed": failed, "batches": batches, "peak_used": peak_used,
            "makespan": makespan}


class GrogServer:
    """Serve ``GridJob`` requests on one HIP device (see the module notes).

    ``hbm_budget`` bytes bound the admitted jobs' estimated footprint,
    ``slots`` the jobs in preparation at once (one worker thread and one
    stream each: the fit's Gram products and the plan compiler are enqueued
    without waiting for them and the same thread solves the fit on the host
    meanwhile, so a preparation's host work overlaps its plan kernels
    without a second thread contending for the interpreter lock),
    ``max_inflight`` the gridding batches on the device.  The worker threads
    live as long as the server (``close()`` or a ``with`` block ends them), so
    successive ``run`` calls reuse their per-thread library state."""

    def __init__(self, device=None, *, mode: str = "fp32", hbm_budget: int | None = None, slots: int = 2,
                 max_inflight: int = 2, keep_results: bool = False):
        if mode not in GRID_MODES:
            raise ValueError(f"unknown gridding mode {mode!r}")
        self.device = _lib.hip_device(device)
        self.mode = mode
        if hbm_budget is None:
            free, total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.5 * free)
        self.hbm_budget = int(hbm_budget)
        self.slots = int(slots)
        self.max_inflight = int(max_inflight)
        self.keep_results = keep_results
        self._streams = [torch.cuda.Stream(device=self.device) for _ in range(self.slots)]
        self._pool = cf.ThreadPoolExecutor(max_workers=self.slots)
        self._grid_stream = torch.cuda.Stream(device=self.device)

    def close(self) -> None:
        """End the worker threads (after the work they hold)."""
        self._pool.shutdown(wait=True)

    def __enter__(self) -> "GrogServer":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    # -- preparation (worker thread, own stream) --------------------------------
    def _fit_finish(self, job: GridJob, fit):
        """The job's operator table from its launched fit (host LAPACK work,
        then the table's assembly on the stream)."""
        return build_table_finish(fit)

    def _prepare(self, job: GridJob, slot: int) -> GridJob:
        st = self._streams[slot]
        # one stream per slot, in this order: the table fit's Gram products
        # (a small kernel, then their copy to the host), the plan compiler,
        # and -- once this thread has solved the fit on the host (LAPACK),
        # which overlaps the plan kernels -- the table, the quantisation and
        # the entry groups; nothing here waits for the device but the Gram
        # copy and the plan's summary
        with torch.cuda.device(self.device), torch.cuda.stream(st):
            fit = build_table_launch(job.calib, PLAN_RADIUS, job.precision, job.lamda, quant=job.quant,
                                     device=self.device)
            pending = GrogInterpolator._launch_plan(job.shape, job.coords, job.oversamp, job.kernel_width,
                                                    job.kernel_shape, device=self.device, defer=True)
            try:
                res = self._fit_finish(job, fit)
            finally:
                pending.wait()  # the plan's buffers and summary stay put until its kernels are done
            g = GrogInterpolator._from_pending(pending)
            g._attach_table(res, job.precision, job.quant, pack=self.mode == "fp32")
            if self.mode != "fp32":
                g._entry_groups(defer=True)
                g._table_bf16(2 if mode_terms(self.mode) == 3 else 1)
            ev = torch.cuda.Event()
            ev.record(st)
        job.interp = g
        job._ready_event = ev
        return job

    # -- device buffers -------------------------------------------------------------
    def _alloc(self, n: int, dtype) -> torch.Tensor:
        """A device buffer of ``n`` elements for a job's gridding (the hook
        a test replaces to inject an allocation failure)."""
        return torch.empty(n, dtype=dtype, device=self.device)

    # -- one gridding batch on the grid stream ---------------------------------
    def _stage_job(self, job: GridJob, st, groups: dict, keep: list) -> None:
        """Transpose (and for the BF16 modes apply) one job on the grid
        stream and add its frames to the batch's reduction groups."""
        g = job.interp
        comp = g.plan._compiled
        y = job.kspace
        if comp.S != 1 or tuple(y.shape) != (comp.C, comp.npts):
            raise ValueError(f"kspace shape {tuple(y.shape)} does not match the plan ({comp.C}, {comp.npts}) "
                             f"of {comp.S} trajectory")
        if y.device != self.device or y.dtype != torch.complex64 or not y.is_contiguous():
            y = y.to(self.device, torch.complex64).contiguous()
        fmt = row_format(self.mode, comp.C)
        yt = self._alloc(max(comp.npts, 1) * sample_row_bytes(comp.C, fmt), torch.uint8)
        out = self._alloc(comp.C * comp.G, torch.complex64).reshape(comp.C, comp.G)
        vbuf = None
        if self.mode != "fp32":
            vbuf = self._alloc(g._rmax() * sample_row_bytes(comp.C, v_format(self.mode)), torch.uint8)
        _lib.call(transpose_entry(fmt, False), _lib.ptr(y), comp.C, comp.npts, _lib.ptr(yt), st.cuda_stream)
        if self.mode != "fp32":
            n_terms = mode_terms(self.mode)
            tbl = g._table_bf16(2 if n_terms == 3 else 1)
            grp = g._entry_groups()[0]
            if grp.n_items > 0:
                _lib.call(apply_entry(fmt), _lib.ptr(yt),
                          _lib.ptr(tbl), _lib.ptr(grp.ent_ptr),
                          _lib.ptr(grp.ent_rep), _lib.ptr(grp.ent_src), _lib.ptr(grp.ent_s),
                          _lib.ptr(grp.items), grp.n_items, int(comp.table.shape[0]), comp.C, n_terms,
                          _lib.ptr(vbuf), st.cuda_stream)
        frames = g._frames(yt, out, 1, vbuf)
        groups.setdefault((comp.C, comp.G), []).extend(frames)
        keep += [y, yt, out, vbuf]
        job.result = out

    def _launch_batch(self, jobs: list) -> tuple:
        """Enqueue one gridding batch; -> (completion event, buffers to keep,
        {job: error} for the jobs that could not be staged)."""
        st = self._grid_stream
        keep = []
        groups = {}
        bad = {}
        with torch.cuda.device(self.device), torch.cuda.stream(st):
            for job in jobs:
                st.wait_event(job._ready_event)
                try:
                    self._stage_job(job, st, groups, keep)
                except (RuntimeError, ValueError, MemoryError) as exc:
                    # nothing of this job reaches the shared launch; what it
                    # already enqueued touches only its own buffers
                    bad[job] = f"launch failed: {type(exc).__name__}: {exc}"
                    job.result = None
            for (C, G), frames in groups.items():
                prog = GridProgram(frames, C, G, self.device, mode_default_tile(self.mode, C),
                                   tile_opts=mode_tile_options(self.mode, C))
                prog.launch(reduce_entry(self.mode))
                keep.append(prog)
            ev = torch.cuda.Event()
            ev.record(st)
        return ev, keep, bad

    # -- the loop ------------------------------------------------------------------
    def run(self, jobs: list, *, timeout: float = 600.0, raw: bool = False) -> dict:
        """Serve ``jobs`` (arrival times relative to the start of the run)
        and return latency / throughput statistics (``raw``: the loop's
        bookkeeping, for ``summarize_runs``).  Rejected and failed jobs are
        listed with their errors and left out of the latency figures."""

        def retire(job):
            if not self.keep_results:
                job.result = None
            job.interp = None

        # garbage-collector pauses during the run, reported with its figures
        pauses, t_gc = [], {}

        def on_gc(phase, info):
            if phase == "start":
                t_gc["t"] = time.perf_counter()
            elif "t" in t_gc:
                pauses.append((info.get("generation", -1), time.perf_counter() - t_gc.pop("t")))

        gc.callbacks.append(on_gc)
        try:
            r = serve_loop(jobs, budget=self.hbm_budget, slots=self.slots, max_inflight=self.max_inflight,
                           estimate=lambda j: estimate_job_bytes(j, self.mode), prepare=self._prepare,
                           launch=self._launch_batch, submit=self._pool.submit, timeout=timeout, on_retire=retire)
        finally:
            gc.callbacks.remove(on_gc)
        r["gc_pauses"] = pauses
        return r if raw else summarize(r, self.mode, self.hbm_budget)


def summarize_runs(runs: list, mode: str, budget: int) -> dict:
    """Statistics of several separate streams pooled: every completed job's
    latency in one distribution, throughput over the summed run times,
    plus each stream's own p99."""
    pooled = {"done": [j for r in runs for j in r["done"]],
              "rejected": [j for r in runs for j in r["rejected"]],
              "failed": [j for r in runs for j in r["failed"]],
              "batches": [b for r in runs for b in r["batches"]],
              "peak_used": max((r["peak_used"] for r in runs), default=0),
              "makespan": sum(r["makespan"] for r in runs),
              "gc_pauses": [p for r in runs for p in r.get("gc_pauses", ())]}
    out = summarize(pooled, mode, budget)
    out["streams"] = len(runs)
    out["p99_ms_per_stream"] = [summarize(r, mode, budget)["p99_ms"] for r in runs]
    return out


def summarize(r: dict, mode: str, budget: int) -> dict:
    """Latency / throughput statistics of a ``serve_loop`` result."""
    done, batches, makespan = r["done"], r["batches"], r["makespan"]
    lat = np.array([j.latency for j in done]) if done else np.zeros(0)
    cs = sum(j.coils * j.samples for j in done)
    # per job kind (2D / 3D): latency and where it goes -- queued until
    # admitted, preparing until the worker thread has enqueued the plan,
    # table and groups (their device work may still run), then until the
    # job's gridding batch completes
    by_kind = {}
    for nd in sorted({len(j.shape) for j in done}):
        js = [j for j in done if len(j.shape) == nd]
        lk = np.array([j.latency for j in js]) * 1e3
        by_kind[f"{nd}D"] = {
            "jobs": len(js), "p50_ms": float(np.percentile(lk, 50)), "p99_ms": float(np.percentile(lk, 99)),
            "queued_ms": float(np.mean([(j.t_admit - j.arrival) * 1e3 for j in js])),
            "prepare_ms": float(np.mean([(j.t_ready - j.t_admit) * 1e3 for j in js])),
            "grid_ms": float(np.mean([(j.t_done - j.t_ready) * 1e3 for j in js])),
        }

    def pct(q):
        return float(np.percentile(lat, q) * 1e3) if lat.size else 0.0

    return {
        "mode": mode, "jobs": len(done), "batches": len(batches),
        "rejected": [(j.job_id, j.error) for j in r["rejected"]],
        "failed": [(j.job_id, j.error) for j in r["failed"]],
        "mean_batch": float(np.mean(batches)) if batches else 0.0,
        "p50_ms": pct(50), "p99_ms": pct(99),
        "mean_ms": float(lat.mean() * 1e3) if lat.size else 0.0, "max_ms": float(lat.max() * 1e3) if lat.size else 0.0,
        "throughput_jobs_s": len(done) / makespan if makespan > 0 else 0.0,
        "coil_samples_per_s": cs / makespan if makespan > 0 else 0.0,
        "makespan_s": makespan, "hbm_budget_gb": budget / 1e9, "peak_admitted_gb": r["peak_used"] / 1e9,
        "by_kind": by_kind,
        "gc_pauses": len(r.get("gc_pauses", ())),
        "gc_pause_ms_max": max((d for _, d in r.get("gc_pauses", ())), default=0.0) * 1e3,
    }
