"""The 3D golden-means kooshball gridding workload at full size against the
oracle: 192^3 grid, 32 coils, 20,000 spokes x 384 samples (7.68 M samples),
2-wide sphere kernel (4 candidate neighbours), shifts quantised to 1/40
(68,921 table entries), lamda = 1e-3, oversampling 1 -- the workload the
gridding throughput is quoted on -- and the same trajectory at 16 coils.

* plan targets and table entries bit-exact against the oracle's float32
  restatement (all 30.72 M candidate replicas);
* the FP32 table against the oracle's FP64 table;
* gridded k-space on a cell subset (the hot k-space centre, where 20,000
  spokes cross, plus random occupied and empty cells): fp32 within 1e-5 of
  the reference flow (src/pygrog/calib/_grog.py:736-902 on the C
  restatement of csrc/torch/grog_interp.cpp:39-73 and
  sparse_ops_cpu_impl.inl:38-120) and of FP64; bf16x3 within 1e-5 of FP64;
  bf16 within 5e-3 of FP64.  At 32 coils the BF16 modes run the grouped
  MFMA apply from 24-bit sample rows (MT=4, KS=2), at 16 coils from
  complex64 rows (MT=2, KS=1).
"""

import numpy as np
import pytest
import torch

from oracle import grid as ogrid
from oracle import plan as oplan
from oracle import table as otable

pytestmark = pytest.mark.gpu

SHAPE = (192, 192, 192)
N_SP, N_RD = 20000, 384


@pytest.fixture(scope="module")
def kooshball_plan():
    from pygrog_amd import synth

    coords = synth.kooshball_3d(N_SP, N_RD)
    p = oplan.attach_fft_plan(oplan.create_plan_single(SHAPE, coords, (1, 1, 1), 2, "sphere"))
    e = oplan.quantize(p["distances"], 41**3, quant=40)
    del p["distances"]
    return coords, p, e


def _cells(p, rng):
    """The 9^3 block around the k-space centre (thousands of replicas per
    cell: the hot pre-pass), 3000 random occupied cells, 200 empty ones."""
    G = int(np.prod(SHAPE))
    c = np.array(SHAPE) // 2
    blk = np.stack(np.meshgrid(*[np.arange(ci - 4, ci + 5) for ci in c], indexing="ij"), -1).reshape(-1, 3)
    centre = np.ravel_multi_index(blk.T, SHAPE)
    occ = np.unique(p["indices"][p["sqrt_weights"] > 0])
    empty = np.setdiff1d(np.arange(G), occ)
    pick = np.concatenate([centre, rng.choice(occ, 3000, replace=False), rng.choice(empty, 200, replace=False)])
    return np.unique(pick)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("C", [32, 16])
def test_kooshball_192_full_size_vs_oracle(device, kooshball_plan, C):
    from pygrog_amd import synth
    from pygrog_amd.calib import GrogInterpolator
    from pygrog_amd.calib._grog import uses_s24

    coords, p, e = kooshball_plan
    print(f"kooshball 192^3 {N_SP}x{N_RD}, C={C}", flush=True)
    g = GrogInterpolator(SHAPE, coords, oversamp=(1, 1, 1), kernel_width=2, kernel_shape="sphere", device=device)
    calib = synth.calibration(SHAPE, C, (20, 20, 20), seed=7, device=device)
    g.calc_interp_table(calib, lamda=1e-3, quant=40, keep_fp64=True)
    assert np.array_equal(g.plan.target_idx.reshape(-1).cpu().numpy(), p["target_idx"].reshape(-1))
    assert np.array_equal(g._interp_idx.reshape(-1).cpu().numpy(), e.reshape(-1))

    t64, ns, nd = otable.table(calib.cpu().numpy(), 1e-3, 1, 40)
    assert (ns, nd) == (41, 3)
    tk = g._interp_kernel.cpu().numpy()
    assert np.abs(tk - t64).max() / np.abs(t64).max() < 1e-6

    y = synth.complex_gaussian((C, N_SP * N_RD), seed=11, device=device)
    yh = y.cpu().numpy()
    cells = _cells(p, np.random.default_rng(C))
    ref = ogrid.grid_reference(yh, p, e, tk, cells=cells)
    g64 = ogrid.grid_fp64(yh, p, e, t64, cells=cells)
    yy = y.reshape(1, 1, C, -1)
    assert uses_s24("bf16x3", C) == (C == 32)
    for mode, against, tol in (("fp32", ref, 1e-5), ("fp32", g64, 1e-5), ("bf16x3", g64, 1e-5),
                               ("bf16x3c", g64, 2e-4), ("bf16", g64, 5e-3)):
        got = g.grid_device(yy, mode).reshape(C, -1)[:, torch.as_tensor(cells, device=device)].cpu().numpy()
        rel = ogrid.rel_l2(got, against)
        print(f"  {mode}: rel-L2 {rel:.2e} (bar {tol:g})", flush=True)
        assert rel <= tol, (mode, rel)


@pytest.mark.timeout(600)
def test_benchmark_fp64_figure_matches_oracle(device, kooshball_plan):
    """The benchmark's own FP64 gridding (the target of its rel-L2 figures)
    equals the oracle's FP64 grid on the same cells, to what two FP64 table
    fits agree on: the device fit solves the ridge normal equations and
    forms the fractional powers from its own eigendecomposition, the
    oracle calls LAPACK gelsd and expm(p logm G); measured 3e-8 at 8 coils,
    against figures of 1e-7 (fp32) and above."""
    import importlib.util
    import pathlib

    from pygrog_amd import synth
    from pygrog_amd.calib import GrogInterpolator

    spec = importlib.util.spec_from_file_location("bench", pathlib.Path(__file__).resolve().parents[1] / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    coords, p, e = kooshball_plan
    C = 8
    g = GrogInterpolator(SHAPE, coords, oversamp=(1, 1, 1), kernel_width=2, kernel_shape="sphere", device=device)
    calib = synth.calibration(SHAPE, C, (20, 20, 20), seed=7, device=device)
    g.calc_interp_table(calib, lamda=1e-3, quant=40, keep_fp64=True)
    y = synth.complex_gaussian((C, N_SP * N_RD), seed=12, device=device)
    cells = _cells(p, np.random.default_rng(3))
    mine = bench.grid_fp64_device(g, y).reshape(C, -1)[:, torch.as_tensor(cells, device=device)].cpu().numpy()
    t64, _, _ = otable.table(calib.cpu().numpy(), 1e-3, 1, 40)
    g64 = ogrid.grid_fp64(y.cpu().numpy(), p, e, t64, cells=cells)
    assert ogrid.rel_l2(mine, g64) <= 1e-7
