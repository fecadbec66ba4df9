"""Wider parity coverage on the HIP device: the 2D radial benchmark case at
full size (256^2, 8 coils, 128 x 512, 3x3 kernel, q = 1/32), coil counts
across every kernel shape (1, 3, 17, 33, 64), SparseFFT scatter with hot
cells (thousands of samples in one cell), and linearity of the grouped
modes on a 3D kooshball."""

import numpy as np
import pytest
import torch

from oracle import grid as ogrid
from oracle import operators as oop
from oracle import plan as oplan
from oracle import table as otable

pytestmark = pytest.mark.gpu


def _radial_full(n_sp, n_rd, seed=0):
    from pygrog_amd import synth

    return synth.radial_2d(n_sp, n_rd, offset=float(np.random.default_rng(seed).uniform(0, np.pi)))


def test_radial_256_8coil_full_size(device):
    """Config-1 sizes: plan and entries bit-exact, FP32 grid within 1e-5 of
    the reference flow, every mode against the FP64 oracle."""
    from pygrog_amd import synth
    from pygrog_amd.calib import GrogInterpolator

    shape, C = (256, 256), 8
    coords = _radial_full(128, 512)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    calib = synth.calibration(shape, C, (24, 24), seed=3, device=device)
    g.calc_interp_table(calib, lamda=1e-3, quant=32, keep_fp64=True)
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, None, 3, "circle"))
    np.testing.assert_array_equal(g.plan.target_idx.reshape(-1).cpu().numpy(), p["target_idx"].reshape(-1))
    e = oplan.quantize(p["distances"], int(g._interp_kernel.shape[0]), quant=32)
    np.testing.assert_array_equal(g._interp_idx.reshape(-1).cpu().numpy(), e.reshape(-1))
    t64, _, _ = otable.table(calib.cpu().numpy(), 1e-3, 1, 32)
    tk = g._interp_kernel.cpu().numpy()
    assert np.abs(tk - t64).max() / np.abs(t64).max() < 1e-6
    y = synth.complex_gaussian((C, 128 * 512), seed=4, device=device)
    ref, _ = ogrid.grid_reference(y.cpu().numpy(), p, e, tk)
    g64 = ogrid.grid_fp64(y.cpu().numpy(), p, e, t64)
    yy = y.reshape(1, 1, C, -1).contiguous()
    got = g.grid_device(yy, "fp32").reshape(C, -1).cpu().numpy()
    assert ogrid.rel_l2(got, ref) <= 1e-5
    assert ogrid.rel_l2(got, g64) <= 1e-5
    assert ogrid.rel_l2(g.grid_device(yy, "bf16x3").reshape(C, -1).cpu().numpy(), g64) <= 1e-5
    assert ogrid.rel_l2(g.grid_device(yy, "bf16").reshape(C, -1).cpu().numpy(), g64) <= 5e-3


@pytest.mark.parametrize("C", [1, 3, 17, 33, 64])
def test_coil_counts_every_kernel_shape(device, C):
    """Every coil group (4, 8, 16, 32, 64 lanes) and MFMA shape, odd and
    even coil counts (unaligned rows), against the oracle."""
    from pygrog_amd.calib import GrogInterpolator

    rng = np.random.default_rng(C)
    shape = (48, 48)
    coords = _radial_full(24, 96, seed=C)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    calib = (rng.standard_normal((C, 16, 16)) + 1j * rng.standard_normal((C, 16, 16))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=1e-2, quant=16, keep_fp64=True)
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, None, 3, "circle"))
    e = g._interp_idx.reshape(p["distances"].shape[:-1]).cpu().numpy()
    y = (rng.standard_normal((C, coords.shape[0] * coords.shape[1]))
         + 1j * rng.standard_normal((C, coords.shape[0] * coords.shape[1]))).astype(np.complex64)
    yy = torch.from_numpy(y).to(device).reshape(1, 1, C, -1).contiguous()
    ref, _ = ogrid.grid_reference(y, p, e, g._interp_kernel.cpu().numpy())
    g64 = ogrid.grid_fp64(y, p, e, g.plan._compiled.table64.cpu().numpy())
    assert ogrid.rel_l2(g.grid_device(yy, "fp32").reshape(C, -1).cpu().numpy(), ref) <= 1e-5
    assert ogrid.rel_l2(g.grid_device(yy, "bf16x3").reshape(C, -1).cpu().numpy(), g64) <= 1e-5
    assert ogrid.rel_l2(g.grid_device(yy, "bf16").reshape(C, -1).cpu().numpy(), g64) <= 5e-3


def test_sparse_fft_scatter_with_hot_cells(device):
    """Thousands of samples in one cell and dozens in its neighbours: the hot
    pre-pass of the SparseFFT scatter against the oracle and float64."""
    from pygrog_amd.operator import SparseFFT

    rng = np.random.default_rng(12)
    grid, image = (24, 24), (20, 20)
    G = 24 * 24
    idx = np.concatenate([rng.integers(0, G, 3000), np.full(2500, 300), rng.integers(290, 310, 1500)])
    rng.shuffle(idx)
    n = idx.size
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    op = SparseFFT(grid, image, idx, w)
    sp = op._sp()
    assert sp.hot[0]["n_hot"] >= 1
    x = (rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))).astype(np.complex64)
    xd = torch.from_numpy(x).to(device)
    g = sp.scatter(xd.reshape(1, 1, 3, n).contiguous())[0, 0]
    ref = torch.zeros(3, G, dtype=torch.complex128, device=device)
    ref.index_add_(1, sp.indices[0], xd[:, sp.sort_perm[0]].to(torch.complex128) * sp.sqrt_w[0].double())
    assert float(torch.linalg.vector_norm(g.to(torch.complex128) - ref) / torch.linalg.vector_norm(ref)) <= 1e-6
    p = np.argsort(idx, kind="stable")
    got = op.adjoint(xd).cpu().numpy()
    assert ogrid.rel_l2(got, oop.sparse_adjoint(x, idx[p], np.sqrt(w)[p], p, grid, image)) <= 1e-5


def test_grouped_modes_linear_on_kooshball(device):
    """Size-independent property on a 3D kooshball: each mode is linear in
    the k-space to its own precision (grid(2 a + b) vs 2 grid(a) + grid(b))."""
    from pygrog_amd import synth
    from pygrog_amd.calib import GrogInterpolator

    shape, C = (48, 48, 48), 16
    coords = synth.kooshball_3d(1500, 96)
    g = GrogInterpolator(shape, coords, oversamp=(1, 1, 1), kernel_width=2, kernel_shape="sphere", device=device)
    g.calc_interp_table(synth.calibration(shape, C, (12, 12, 12), seed=5, device=device), lamda=1e-3, quant=20)
    a = synth.complex_gaussian((1, 1, C, 1500 * 96), seed=6, device=device)
    b = synth.complex_gaussian((1, 1, C, 1500 * 96), seed=7, device=device)
    for mode, tol in (("fp32", 1e-6), ("bf16x3", 1e-5), ("bf16", 5e-3)):
        lhs = g.grid_device((2 * a + b).contiguous(), mode).clone()
        rhs = 2 * g.grid_device(a, mode).clone() + g.grid_device(b, mode)
        rel = float(torch.linalg.vector_norm(lhs - rhs) / torch.linalg.vector_norm(rhs))
        assert rel <= tol, (mode, rel)


@pytest.mark.parametrize("C", [3, 16, 33, 64])
@pytest.mark.parametrize("hot", [True, False])
def test_every_tile_shape_same_bits(device, monkeypatch, C, hot):
    """Every compiled launch shape -- the wave-tile kernel (256 cells per
    workgroup, one wavefront per 64 cells) and the workgroup tiles -- at two
    persistent grid sizes gives the default shape's grid bit for bit, in
    every mode, with the hot pre-pass and with the multi-unit fold (hot
    threshold raised); the fp32 grid also matches the oracle."""
    import pygrog_amd.calib._grog as grog_mod
    from pygrog_amd import _lib
    from pygrog_amd.calib import GrogInterpolator

    if not hot:
        monkeypatch.setattr(grog_mod, "HOT_MIN_UNITS", 10**9)
    rng = np.random.default_rng(100 + C)
    shape = (80, 80)
    coords = _radial_full(40, 160, seed=C)  # DC and the first rings hold many replicas
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    calib = (rng.standard_normal((C, 16, 16)) + 1j * rng.standard_normal((C, 16, 16))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=1e-2, quant=16)
    assert (g.plan._compiled.hot[0]["n_hot"] > 0) == hot
    n = coords.shape[0] * coords.shape[1]
    y = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    yy = torch.from_numpy(y).to(device).reshape(1, 1, C, -1).contiguous()
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, None, 3, "circle"))
    e = g._interp_idx.reshape(p["distances"].shape[:-1]).cpu().numpy()
    ref, _ = ogrid.grid_reference(y, p, e, g._interp_kernel.cpu().numpy())
    for mode in ("fp32", "bf16x3", "bf16"):
        base = g.grid_device(yy, mode).clone()
        if mode == "fp32":
            assert ogrid.rel_l2(base.reshape(C, -1).cpu().numpy(), ref) <= 1e-5
        for tc in _lib.tile_options(C):
            for nb in (8, None):
                g._tuned = {mode: (tc, nb)}
                got = g.grid_device(yy, mode, program_cache={})
                assert torch.equal(got, base), (mode, tc, nb)
        g._tuned = {}


@pytest.mark.parametrize("C", [3, 16, 40])
def test_scatter_every_tile_shape_same_bits(device, monkeypatch, C):
    """The SparseFFT scatter (identity operator, coil-major sources) over
    every compiled launch shape, with a hot cell: identical bits."""
    from pygrog_amd import _lib
    from pygrog_amd.operator import SparseFFT

    rng = np.random.default_rng(40 + C)
    grid, image = (32, 32), (28, 28)
    G = 32 * 32
    idx = np.concatenate([rng.integers(0, G, 4000), np.full(900, 500), rng.integers(480, 520, 700)])
    rng.shuffle(idx)
    n = idx.size
    w = rng.uniform(0.5, 2.0, n).astype(np.float32)
    sp = SparseFFT(grid, image, idx, w)._sp()
    assert sp.hot[0]["n_hot"] >= 1
    x = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    xd = torch.from_numpy(x).to(device).reshape(1, 1, C, n).contiguous()
    base = sp.scatter(xd).clone()
    ref = torch.zeros(C, G, dtype=torch.complex128, device=device)
    ref.index_add_(1, sp.indices[0], xd[0, 0][:, sp.sort_perm[0]].to(torch.complex128) * sp.sqrt_w[0].double())
    rel = torch.linalg.vector_norm(base[0, 0].to(torch.complex128) - ref) / torch.linalg.vector_norm(ref)
    assert float(rel) <= 1e-6
    lib = _lib.load()
    for tc in _lib.tile_options(C):
        monkeypatch.setattr(lib, "pg_grid_tile_cells", lambda c, tc=tc: tc)
        assert torch.equal(sp.scatter(xd), base), tc


@pytest.mark.parametrize("C", [8, 16])
def test_long_units_every_tile_shape(device, monkeypatch, C):
    """Reduction units longer than a wave tile's row batch and than the
    parallel hot pre-pass holds (unit chunk 80 instead of 32): the wave-tile
    kernels fold such cells from memory and the pre-pass falls back to the
    serial fold; every launch shape gives the same bits and the fp32 grid
    matches the oracle."""
    import pygrog_amd.calib._grog as grog_mod
    from pygrog_amd import _lib
    from pygrog_amd.calib import GrogInterpolator

    monkeypatch.setattr(grog_mod, "UNIT_CHUNK", 80)
    rng = np.random.default_rng(7 + C)
    shape = (64, 64)
    coords = _radial_full(96, 128, seed=C)  # cells near the centre hold 40-300 replicas
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    calib = (rng.standard_normal((C, 16, 16)) + 1j * rng.standard_normal((C, 16, 16))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=1e-2, quant=16)
    cp = g.plan._compiled.cell_ptr[0].to(torch.int64)
    assert int((cp[1:] - cp[:-1]).max()) > 80  # some cells span several long units
    n = coords.shape[0] * coords.shape[1]
    y = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    yy = torch.from_numpy(y).to(device).reshape(1, 1, C, -1).contiguous()
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, None, 3, "circle"))
    e = g._interp_idx.reshape(p["distances"].shape[:-1]).cpu().numpy()
    ref, _ = ogrid.grid_reference(y, p, e, g._interp_kernel.cpu().numpy())
    for mode in ("fp32", "bf16x3"):
        base = g.grid_device(yy, mode).clone()
        if mode == "fp32":
            assert ogrid.rel_l2(base.reshape(C, -1).cpu().numpy(), ref) <= 1e-5
        for tc in _lib.tile_options(C):
            g._tuned = {mode: (tc, None)}
            assert torch.equal(g.grid_device(yy, mode, program_cache={}), base), (mode, tc)
        g._tuned = {}
