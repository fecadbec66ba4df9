"""GROG gridding benchmark on one MI355X.

Workload (default ``kooshball3d``): 3D golden-means kooshball on a 192^3 grid,
32 coils, 20,000 spokes x 384 samples (7.68 M samples), complex64 k-space
i.i.d. complex Gaussian (sigma 1 per component), operator table fit from a
20^3 calibration region of a seeded ellipsoid phantom times 32 Gaussian coil
maps, shifts quantised to 1/40 grid step (41^3 = 68,921 table entries),
2-wide sphere kernel (4 candidate neighbours), lamda = 1e-3, oversampling 1.

A step grids the device-resident k-space (1, 32, 7.68 M) into a dense
(32, 192^3) complex64 grid with the compiled plan and table already resident
(``GrogInterpolator.grid_device``).  Modes:
  fp32    plain fused kernel (per-replica operator apply, FP32);
  bf16    grouped complex GEMM on BF16 MFMA, FP32 accumulate (1 product);
  bf16x3  grouped, split-BF16 (3 products, ~FP32 accuracy).
``value`` is the fastest BF16-MFMA mode whose rel-L2 vs the FP64 reference is
<= 1e-3, in coil-samples/s = samples * coils / step time.

The JSON line also carries: every mode's throughput and rel-L2, the roofline
of the headline mode's dominant kernel (algorithmic bytes / HIP-event kernel
time vs 8 TB/s), and the CPU baseline: the reference's own CPU gridding flow
(``interpolate(grid=True)``: csrc/torch ops compiled in ``oracle/_ref``) on a
bounded spoke sample, timed on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ...]
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np
import torch

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "coil-samples/s gridded, BF16-MFMA vs FP32 @relL2≤1e-3; p99 serve latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def make_workload(name: str, dev):
    from pygrog_amd import synth
    from pygrog_amd.calib import GrogInterpolator

    if name == "kooshball3d":
        shape, C, n_sp, n_rd = (192, 192, 192), 32, 20000, 384
        coords = synth.kooshball_3d(n_sp, n_rd)
        g = GrogInterpolator(shape, coords, oversamp=(1, 1, 1), kernel_width=2, kernel_shape="sphere", device=dev)
        calib = synth.calibration(shape, C, (20, 20, 20), seed=7, device=dev)
        g.calc_interp_table(calib, lamda=1e-3, quant=40, keep_fp64=True)
        desc = ("3D golden-means kooshball 192^3, 32 coils, 20000 spokes x 384 samples (7.68M), "
                "q=1/40 (68921-entry table), 2-wide sphere kernel, lamda=1e-3, oversamp 1")
        y = synth.complex_gaussian((1, 1, C, n_sp * n_rd), seed=11, device=dev)
        spokes = n_sp
    elif name == "radial2d":
        shape, C, n_sp, n_rd = (256, 256), 8, 128, 512
        coords = synth.radial_2d(n_sp, n_rd)
        g = GrogInterpolator(shape, coords, kernel_width=3, device=dev)
        calib = synth.calibration(shape, C, (24, 24), seed=3, device=dev)
        g.calc_interp_table(calib, lamda=1e-3, quant=32, keep_fp64=True)
        desc = "2D golden-angle radial 256^2, 8 coils, 128x512, q=1/32, 3x3 kernel, lamda=1e-3"
        y = synth.complex_gaussian((1, 1, C, n_sp * n_rd), seed=11, device=dev)
        spokes = n_sp
    else:
        raise SystemExit(f"unknown workload {name}")
    return {"name": name, "g": g, "y": y, "C": C, "N": y.shape[-1], "shape": shape, "desc": desc,
            "spokes": spokes, "n_read": n_rd, "coords": coords}


# ---------------------------------------------------------------------------
# FP64 reference on the device (rel-L2 of every mode)
# ---------------------------------------------------------------------------
def grid_fp64_device(g, y: torch.Tensor, chunk: int = 1 << 16) -> torch.Tensor:
    """sum_r w_r K64[e_r] y[:, src_r] in float64 (y: (C, N))."""
    comp = g.plan._compiled
    T = comp.table64
    C, N = y.shape
    t = comp.target_flat.reshape(-1)
    e = comp.interp_idx.reshape(-1)
    w = g.plan._dev["weights"].reshape(-1)
    valid = torch.nonzero(t >= 0).squeeze(1)
    out = torch.zeros(C, comp.G, 2, dtype=torch.float64, device=y.device)
    y64 = y.to(torch.complex128)
    for i in range(0, valid.numel(), chunk):
        v = valid[i:i + chunk]
        Kg = T[e[v]].contiguous()
        Y = y64[:, v // comp.kw].T.contiguous().unsqueeze(-1)
        V = torch.bmm(Kg, Y).squeeze(-1) * w[v].to(torch.float64)[:, None]
        out.index_add_(1, t[v], torch.view_as_real(V.T.contiguous()))
    return torch.view_as_complex(out)


def rel_l2(a: torch.Tensor, b: torch.Tensor) -> float:
    a = a.to(torch.complex128)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


# ---------------------------------------------------------------------------
# algorithmic bytes per kernel launch (the roofline numerators)
# ---------------------------------------------------------------------------
def alg_bytes(kernel: str, st: dict) -> float:
    C, N, G, R, U, K = st["C"], st["N"], st["G"], st["R"], st["U"], st["K"]
    if kernel == "pg_grid_fp32":
        # source rows of every replica + grid write + replica records (src,
        # entry, sqrt w) + units + cell map + table once
        return 8 * C * R + 8 * C * G + 12 * R + 8 * U + 4 * G + 8 * C * C * K
    if kernel == "pg_apply_grouped_bf16":
        parts = st["n_parts"]
        # y_t rows of every replica + table (bf16, real 2C x 2C per entry) + V write + records
        return 8 * C * R + 2 * parts * 4 * C * C * K + 8 * C * R + 12 * R + 4 * K
    if kernel == "pg_grid_reduce_v":
        return 8 * C * R + 8 * C * G + 8 * U + 4 * G
    if kernel == "pg_transpose_cn_c64":
        return 16 * C * N
    raise KeyError(kernel)


# ---------------------------------------------------------------------------
# HBM traffic from rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; KiB/dispatch)
# ---------------------------------------------------------------------------
# kernels launched by one call of each entry point; k_hot_cells is shared by
# the fp32 and reduce entries (same plan, same hot cells)
ENTRY_KERNELS = {
    "pg_grid_fp32": [r"k_hot_units<\d+, 0>", r"k_hot_cells<", r"k_grid<\d+, 0>"],
    "pg_grid_reduce_v": [r"k_hot_units<\d+, 1>", r"k_hot_cells<", r"k_grid<\d+, 1>"],
    "pg_apply_grouped_bf16:1": [r"k_apply_grouped<\d+, \d+, 1>"],
    "pg_apply_grouped_bf16:3": [r"k_apply_grouped<\d+, \d+, 3>"],
    "pg_transpose_cn_c64": [r"k_transpose_cn"],
}
COPY_KERNEL = r"neg_kernel"


def _pmc_dispatches(path: str, counter: str) -> list:
    """[(dispatch id, kernel name, value)] of ``counter``, in dispatch order."""
    import csv
    import re

    per = {}
    names = {}
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            d = int(r["Dispatch_Id"])
            per[d] = per.get(d, 0.0) + float(r["Counter_Value"])
            names[d] = re.sub(r"\(.*$", "", r["Kernel_Name"]).replace("void ", "")
    return [(d, names[d], per[d]) for d in sorted(per)]


def _pmc_means(rows: list) -> dict:
    out = {}
    for _, name, v in rows:
        out.setdefault(name, []).append(v)
    return {k: float(np.mean(v)) for k, v in out.items()}


def summarize_pmc(fetch_csv: str, write_csv: str, copy_bytes: int, workload: str) -> dict:
    import re

    fetch_rows = _pmc_dispatches(fetch_csv, "FETCH_SIZE")
    write_rows = _pmc_dispatches(write_csv, "WRITE_SIZE")
    fetch, write = _pmc_means(fetch_rows), _pmc_means(write_rows)

    def copy_kib(rows):
        # the calibration copy is the first negation the benchmark dispatches
        return next(v for _, n, v in rows if re.search(COPY_KERNEL, n))

    def pick(table, pat):
        hits = [v for k, v in table.items() if re.search(pat, k)]
        if len(hits) != 1:
            raise SystemExit(f"pattern {pat!r} matches {len(hits)} kernels")
        return hits[0]

    ff = copy_bytes / (copy_kib(fetch_rows) * 1024.0)
    wf = copy_bytes / (copy_kib(write_rows) * 1024.0)
    entries = {}
    for entry, pats in ENTRY_KERNELS.items():
        try:
            fb = sum(pick(fetch, p) for p in pats) * 1024.0 * ff
            wb = sum(pick(write, p) for p in pats) * 1024.0 * wf
        except SystemExit:
            continue
        entries[entry] = {"fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb}
    return {"workload": workload, "copy_bytes": copy_bytes, "fetch_factor": ff, "write_factor": wf,
            "entries": entries,
            "kernels_kib": {k: {"FETCH_SIZE": fetch.get(k), "WRITE_SIZE": write.get(k)} for k in sorted(fetch)
                            if k.startswith("pg::")}}


# ---------------------------------------------------------------------------
# reference CPU baseline (bounded sample)
# ---------------------------------------------------------------------------
def cpu_baseline(wl, n_spokes: int, threads: int) -> dict:
    """Time the reference's CPU flow of interpolate(grid=True)
    (src/pygrog/calib/_grog.py:668-695, 736-821 on csrc/torch CPU ops)
    on the first ``n_spokes`` spokes."""
    import oracle

    ref = oracle.ref_ext()
    kind = "reference"
    if ref is None:
        kind = "port"
    g = wl["g"]
    comp = g.plan._compiled
    kw, C = comp.kw, wl["C"]
    n = n_spokes * wl["n_read"]
    nper = n * kw
    tgt = comp.target_flat.reshape(-1)[:nper].cpu()
    w = g.plan._dev["weights"].reshape(-1)[:nper].cpu()
    e = comp.interp_idx.reshape(-1)[:nper].cpu()
    kernel = comp.table.cpu()
    data = wl["y"][0, 0, :, :n].cpu().contiguous()
    # the sample's FFT plan, as _attach_fft_plan builds it (untimed setup)
    idx = tgt.clone()
    sqrt_w = torch.sqrt(w.clamp(min=0.0))
    bad = idx < 0
    idx[bad] = 0
    sqrt_w[bad] = 0.0
    sort_perm = torch.argsort(idx)
    sidx = idx[sort_perm].contiguous()
    sqw = sqrt_w[sort_perm].contiguous()
    inv_perm = torch.empty_like(sort_perm)
    inv_perm[sort_perm] = torch.arange(nper)
    pre_w = sqw[inv_perm]
    G = comp.G
    torch.set_num_threads(threads)

    def run():
        # _sparse_full_single: replicate, transpose, apply, permute back
        data_rep = data.reshape(C, n).unsqueeze(-1).expand(-1, -1, kw).reshape(C, n * kw)
        data_interp = data_rep.T.reshape(n * kw, 1, C).contiguous()
        if ref is not None:
            res = ref.grog_interpolate(data_interp, e, kernel)
        else:
            from oracle import ops

            res = torch.from_numpy(ops.interpolate(data_interp.numpy(), e.numpy(), kernel.numpy()))
        data_interp.copy_(res)
        sparse = data_interp.permute(1, 2, 0).contiguous().reshape(C, nper)
        # _grid_kspace: pre-weights, density, per-coil sort + scatter
        sparse_dc = sparse * pre_w.to(sparse.dtype).unsqueeze(0)
        grid = torch.zeros(C, G, dtype=torch.complex64)
        dens = torch.zeros(G, dtype=torch.float32)
        dens.index_add_(0, sidx, (sqw * sqw).contiguous())
        for c in range(C):
            sc = sparse_dc[c][sort_perm]
            if ref is not None:
                ref.scatter_add(grid[c], sc, sidx, sqw)
            else:
                from oracle import ops

                gc_ = grid[c].numpy()
                ops.scatter_add(gc_, sc.numpy(), sidx.numpy(), sqw.numpy())
        return grid

    t0 = time.perf_counter()
    run()
    t1 = time.perf_counter()
    return {"seconds": t1 - t0, "coil_samples": n * C, "kind": kind, "spokes": n_spokes}


# ---------------------------------------------------------------------------
# batched dynamic frames: 200 x 2D radial 192^2, 16 coils, one shared table
# ---------------------------------------------------------------------------
FRAMES_DESC = ("200 dynamic 2D golden-angle radial frames on 192^2, 16 coils, 512 samples/spoke, spokes per frame "
               "U{13..55} (seeded), seeded golden-angle rotation per frame, one table from a 24^2 ACR "
               "(q=1/32, 3x3 kernel, lamda=1e-3), all frames in one HIP-graph launch")


def make_frames2d(dev, n_frames: int = 200, seed: int = 2024):
    from pygrog_amd import synth
    from pygrog_amd.calib import GrogFrameBatch

    shape, C, n_rd = (192, 192), 16, 512
    rng = np.random.default_rng(seed)
    spokes = rng.integers(13, 56, size=n_frames)
    # frame f continues one golden-angle stream from a seeded starting spoke
    starts = rng.integers(0, 10000, size=n_frames)
    golden = np.deg2rad(synth.GOLDEN_ANGLE_DEG)
    coords = [synth.radial_2d(int(n), n_rd, offset=float(golden * s)) for n, s in zip(spokes, starts)]
    b = GrogFrameBatch(shape, coords, kernel_width=3, device=dev)
    calib = synth.calibration(shape, C, (24, 24), seed=seed + 1, device=dev)
    b.calc_interp_table(calib, lamda=1e-3, quant=32, keep_fp64=True)
    ys = [synth.complex_gaussian((C, int(n) * n_rd), seed=seed + 100 + f, device=dev) for f, n in enumerate(spokes)]
    return b, b.pack(ys), spokes


def frames_fp64_rel(b, y, out) -> float:
    num = den = 0.0
    for f, fr in enumerate(b.frames):
        C, N = b.C, int(b.n_samples[f])
        yf = y[C * int(b.sample_off[f]):C * int(b.sample_off[f + 1])].reshape(C, N)
        ref = grid_fp64_device(fr, yf)
        num += float(torch.linalg.vector_norm(out[f].reshape(C, -1).to(torch.complex128) - ref) ** 2)
        den += float(torch.linalg.vector_norm(ref) ** 2)
    return (num / den) ** 0.5


def run_frames2d(args, dev) -> dict:
    from pygrog_amd.calib._grog import KernelTimer

    t0 = time.perf_counter()
    b, y, spokes = make_frames2d(dev)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    st = b.work_stats()
    log(f"frames2d setup {setup_s:.1f}s: {st}")
    C, N, G, F = st["C"], st["samples"], st["G"], st["F"]
    stats = {"C": C, "N": N, "G": G * F, "R": st["replicas_valid"], "U": st["units"], "K": st["table_entries"]}
    modes = {}
    for mode in [m for m in args.modes.split(",") if m]:
        out = b.new_output()
        tune = b.autotune(y, mode, out)
        graph, out = b.capture(y, mode, out)
        for _ in range(args.warmup):
            graph.replay()
        torch.cuda.synchronize()
        rl = None if args.no_check else frames_fp64_rel(b, y, out)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            graph.replay()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        # per-kernel durations: the same launches, eager, bracketed by HIP events
        timer = KernelTimer(dev)
        for _ in range(max(3, min(args.steps, 10))):
            b.grid(y, mode, out, timer=timer)
        torch.cuda.synchronize()
        kt = {k: float(np.mean(v)) for k, v in timer.durations().items()}
        modes[mode] = {"coil_samples_per_s": N * C / (ms * 1e-3), "ms_per_step": ms, "rel_l2_vs_fp64": rl,
                       "kernel_ms": kt, "autotune": {"tile_cells": tune["tile_cells"], "n_blocks": tune["n_blocks"]}}
        log(f"frames2d {mode}: {ms:.3f} ms/step (graph), {N * C / (ms * 1e-3):.3e} coil-samples/s, rel-L2 {rl}, "
            f"kernels {kt}, tune {tune}")
    all_roof = {}
    for m, v in modes.items():
        stats["n_parts"] = 2 if m == "bf16x3" else 1
        all_roof[m] = {}
        for k, t in v["kernel_ms"].items():
            kk = "pg_transpose_cn_c64" if k == "pg_transpose_frames_c64" else k
            all_roof[m][k] = {"GB/s": alg_bytes(kk, stats) / (t * 1e-3) / 1e9, "ms": t}
    head = "fp32" if "fp32" in modes else next(iter(modes))
    kname = "pg_grid_fp32" if head == "fp32" else max(modes[head]["kernel_ms"], key=modes[head]["kernel_ms"].get)
    kms = modes[head]["kernel_ms"][kname]
    stats["n_parts"] = 1
    ab = alg_bytes(kname, stats)
    roof = {"kernel": kname, "bound": "hbm", "achieved": ab / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ab / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None, "alg_bytes": ab,
            "kernel_ms": kms}
    traffic_file = ROOT / "profiles" / "traffic_frames2d.json"
    if traffic_file.exists():
        tr = json.loads(traffic_file.read_text()).get("entries", {})
        if kname in tr:
            roof["traffic"] = tr[kname]["traffic_bytes"]
    bf = [m for m in modes if m.startswith("bf16") and (modes[m]["rel_l2_vs_fp64"] is None
                                                       or modes[m]["rel_l2_vs_fp64"] <= 1e-3)]
    best = max(bf, key=lambda m: modes[m]["coil_samples_per_s"]) if bf else head
    hm = modes[best]
    return {
        "metric": METRIC, "value": hm["coil_samples_per_s"], "unit": "coil-samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": hm["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16" if best.startswith("bf16") else "fp32",
        "data": "synthetic (seeded trajectories and phantom; Gaussian k-space)",
        "config": {"workload": FRAMES_DESC, "mode": best, "frames": F, "samples": N, "coils": C,
                   "grid": list(b.grid_shape), "spokes_min_max": [int(spokes.min()), int(spokes.max())],
                   "table_entries": stats["K"], "replicas_valid": stats["R"]},
        "modes": modes, "roofline": roof, "roofline_all": all_roof, "cpu_baseline": None, "setup_s": setup_s,
    }


# ---------------------------------------------------------------------------
def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="kooshball3d", choices=["kooshball3d", "radial2d", "frames2d"])
    ap.add_argument("--modes", default="fp32,bf16,bf16x3")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline work (s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 rel-L2 check")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="run one plain device copy (float32 negation) of BYTES first: the PMC scale factor")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="turn two rocprofv3 --pmc counter_collection.csv files into profiles/traffic_<workload>.json "
                         "(no GPU needed)")
    args = ap.parse_args(argv)

    if args.summarize_pmc:
        if args.copy_calibration <= 0:
            raise SystemExit("--summarize-pmc needs --copy-calibration BYTES of the profiled run")
        res = summarize_pmc(args.summarize_pmc[0], args.summarize_pmc[1], args.copy_calibration, args.workload)
        path = ROOT / "profiles" / f"traffic_{args.workload}.json"
        path.write_text(json.dumps(res, indent=1) + "\n")
        print(json.dumps(res["entries"]))
        return 0

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus N>1 is not supported (see DESIGN.md, "
                                   "'How the path would shard')", "n_gpus": args.gpus}))
        return 2
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        return 2
    from pygrog_amd.calib._grog import KernelTimer

    dev = torch.device("cuda:0")
    if args.workload == "frames2d":
        line = run_frames2d(args, dev)
        print(json.dumps(line))
        return 0
    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    t_setup = time.perf_counter()
    wl = make_workload(args.workload, dev)
    g = wl["g"]
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    log(f"setup {setup_s:.1f}s: {g.work_stats()}")
    st = g.work_stats()
    comp = g.plan._compiled
    stats = {"C": wl["C"], "N": wl["N"], "G": comp.G, "R": st["replicas_valid"], "U": st["units"],
             "K": st["table_entries"]}
    y = wl["y"]
    out = torch.empty(1, 1, wl["C"], comp.G, dtype=torch.complex64, device=dev)
    ref64 = None
    if not args.no_check:
        ref64 = grid_fp64_device(g, y[0, 0])
        torch.cuda.synchronize()

    modes = {}
    cache = {}
    for mode in [m for m in args.modes.split(",") if m]:
        for _ in range(args.warmup):
            g.grid_device(y, mode, out, program_cache=cache)
        torch.cuda.synchronize()
        rl = rel_l2(out[0, 0], ref64) if ref64 is not None else None
        timer = KernelTimer(dev)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            g.grid_device(y, mode, out, timer=timer, program_cache=cache)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        kt = {k: float(np.mean(v)) for k, v in timer.durations().items()}
        modes[mode] = {"coil_samples_per_s": wl["N"] * wl["C"] / (ms * 1e-3), "ms_per_step": ms,
                       "rel_l2_vs_fp64": rl, "kernel_ms": kt}
        log(f"{mode}: {ms:.3f} ms/step, {modes[mode]['coil_samples_per_s']:.3e} coil-samples/s, rel-L2 {rl}, "
            f"kernels {kt}")

    bf = [m for m in modes if m.startswith("bf16") and (modes[m]["rel_l2_vs_fp64"] is None
                                                       or modes[m]["rel_l2_vs_fp64"] <= 1e-3)]
    head = max(bf, key=lambda m: modes[m]["coil_samples_per_s"]) if bf else next(iter(modes))
    hm = modes[head]
    kname, kms = max(hm["kernel_ms"].items(), key=lambda kv: kv[1])
    stats["n_parts"] = 2 if head == "bf16x3" else 1
    ab = alg_bytes(kname, stats)
    achieved = ab / (kms * 1e-3) / 1e9
    roof = {"kernel": kname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "alg_bytes": ab, "kernel_ms": kms}
    traffic_file = ROOT / "profiles" / f"traffic_{args.workload}.json"
    tkey = kname + (f":{3 if head == 'bf16x3' else 1}" if kname == "pg_apply_grouped_bf16" else "")
    if traffic_file.exists():
        tr = json.loads(traffic_file.read_text()).get("entries", {})
        if tkey in tr:
            roof["traffic"] = tr[tkey]["traffic_bytes"]
            roof["traffic_source"] = f"profiles/{traffic_file.name} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)"
    all_roof = {}
    for m, v in modes.items():
        stats["n_parts"] = 2 if m == "bf16x3" else 1
        all_roof[m] = {k: {"GB/s": alg_bytes(k, stats) / (t * 1e-3) / 1e9, "ms": t} for k, t in v["kernel_ms"].items()}

    cpu = None
    if not args.no_cpu_baseline:
        threads = len(os.sched_getaffinity(0))
        probe = cpu_baseline(wl, max(1, min(wl["spokes"], 100 if wl["spokes"] > 1000 else wl["spokes"])), threads)
        rate = probe["coil_samples"] / probe["seconds"]
        n_sp = int(min(wl["spokes"], max(1, args.cpu_seconds * rate / (wl["n_read"] * wl["C"]))))
        res = cpu_baseline(wl, n_sp, threads)
        cpu = {"value": res["coil_samples"] / res["seconds"], "unit": "coil-samples/s", "cores": threads,
               "kind": res["kind"],
               "sample": f"first {n_sp} of {wl['spokes']} spokes ({n_sp * wl['n_read']} samples x {wl['C']} coils), "
                         f"reference interpolate(grid=True) flow, {res['seconds']:.2f} s"}
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC,
        "value": hm["coil_samples_per_s"],
        "unit": "coil-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if head.startswith("bf16") else "fp32",
        "data": "synthetic (seeded trajectory, phantom, coil maps; Gaussian k-space)",
        "config": {"workload": wl["desc"], "mode": head, "samples": wl["N"], "coils": wl["C"],
                   "grid": list(wl["shape"]), "table_entries": stats["K"], "replicas_valid": stats["R"]},
        "modes": modes,
        "speedup_bf16_vs_fp32": (hm["coil_samples_per_s"] / modes["fp32"]["coil_samples_per_s"]
                                 if "fp32" in modes else None),
        "roofline": roof,
        "roofline_all": all_roof,
        "cpu_baseline": cpu,
        "setup_s": setup_s,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
