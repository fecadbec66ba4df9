"""GPU plan compiler and fused gridding vs the oracle.

* target indices, shifts, quantised table entries, sorted indices and
  permutations: bit-exact vs the numpy float32 restatement;
* EGROG weights: 1e-6 relative (float32 exp differs in the last ulp);
* density: bit-exact (same stable summation order);
* gridded k-space (FP32): rel-L2 <= 1e-5 vs the reference flow on the C
  oracle with the same table, and reported vs the FP64 oracle;
* BF16 grouped modes: rel-L2 vs the FP64 oracle <= 5e-3 (1 product) and
  <= 1e-5 (3 split products);
* determinism: bitwise identical across runs and alone vs batched.
"""

import numpy as np
import pytest
import torch

from oracle import grid as ogrid
from oracle import plan as oplan
from oracle import table as otable
from pygrog_amd.calib import GrogInterpolator

pytestmark = pytest.mark.gpu


def _radial(n_spokes, n_read, seed=0):
    rng = np.random.default_rng(seed)
    ga = np.deg2rad(111.246117975)
    r = (np.arange(n_read) - n_read // 2) / n_read
    th = ga * np.arange(n_spokes) + rng.uniform(0, np.pi)
    return np.stack([np.outer(np.cos(th), r), np.outer(np.sin(th), r)], -1).astype(np.float32)


def _calib(C, shape, seed=1):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((C, *shape)) + 1j * rng.standard_normal((C, *shape))).astype(np.complex64)


CASES = [
    # name, shape, coords-maker, oversamp, width, kshape, C, calib shape, quant/precision
    ("2d_rand_osf2", (12, 12), lambda: np.random.default_rng(5).standard_normal((5, 7, 2)).astype(np.float32), 2.0, 2,
     "circle", 3, (12, 12), dict(precision=1)),
    ("2d_radial_w3", (64, 64), lambda: _radial(32, 128), None, 3, "circle", 8, (24, 24), dict(quant=32)),
    ("2d_square_osf15", (40, 40), lambda: _radial(20, 80, 3), 1.5, 2, "square", 4, (16, 16), dict(precision=1)),
    ("2d_cross_w3", (32, 32), lambda: _radial(16, 64, 4), 2.0, 3, "cross", 5, (16, 16), dict(precision=2)),
    ("3d_sphere", (16, 16, 16), lambda: np.random.default_rng(7).uniform(-0.5, 0.5, (30, 16, 3)).astype(np.float32),
     (1, 1, 1), 2, "sphere", 6, (8, 8, 8), dict(quant=10)),
    ("3d_default_osf", (12, 12, 12), lambda: np.random.default_rng(8).uniform(-0.5, 0.5, (40, 8, 3)).astype(np.float32),
     None, 2, "cube", 4, (6, 6, 6), dict(precision=1)),
]


def _build(case, device):
    name, shape, mk, osf, w, ks, C, cshape, q = case
    coords = mk()
    g = GrogInterpolator(shape, coords, oversamp=osf, kernel_width=w, kernel_shape=ks, device=device)
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, osf, w, ks))
    return g, p, coords


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_plan_bit_exact(device, case):
    g, p, _ = _build(case, device)
    assert tuple(g.plan.grid_shape) == p["grid_shape"]
    np.testing.assert_array_equal(g.plan.target_idx.reshape(-1).cpu().numpy(), p["target_idx"].reshape(-1))
    d_gpu = g.plan.distances.cpu().numpy().reshape(p["distances"].shape)
    assert np.array_equal(d_gpu.view(np.uint32), p["distances"].view(np.uint32)), "shifts not bit-exact"
    np.testing.assert_array_equal(g.plan.indices.cpu().numpy(), p["indices"])
    np.testing.assert_array_equal(g.plan.sort_perm.cpu().numpy(), p["sort_perm"])
    np.testing.assert_array_equal(g.plan.inv_perm.cpu().numpy(), p["inv_perm"])
    np.testing.assert_allclose(g.plan.weights.reshape(-1).cpu().numpy(), p["weights"].reshape(-1), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(g.plan.sqrt_weights.cpu().numpy(), p["sqrt_weights"], rtol=1e-6, atol=1e-7)
    assert g.plan.n_samples == p["n_samples"]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_quantize_bit_exact_and_grid_parity(device, case):
    name, shape, mk, osf, w, ks, C, cshape, q = case
    g, p, _ = _build(case, device)
    calib = _calib(C, cshape)
    g.calc_interp_table(calib, lamda=0.01, **q)
    table = g._interp_kernel.cpu().numpy()
    e_ref = oplan.quantize(p["distances"], table.shape[0], precision=q.get("precision", 1), quant=q.get("quant"))
    np.testing.assert_array_equal(g._interp_idx.reshape(-1).cpu().numpy(), e_ref.reshape(-1))
    # table vs FP64 oracle (same algorithm, numpy/scipy)
    t64, ns, nd = otable.table(calib, 0.01, q.get("precision", 1), q.get("quant"))
    assert np.abs(table - t64).max() / np.abs(t64).max() < 1e-6
    # gridding vs the reference flow (same complex64 table)
    rng = np.random.default_rng(11)
    npts = p["target_idx"].shape[:-1]
    y = (rng.standard_normal((C, *npts)) + 1j * rng.standard_normal((C, *npts))).astype(np.complex64)
    kgrid, mplan = g.interpolate(y, grid=True)
    gref, dens = ogrid.grid_reference(y.reshape(C, -1), p, e_ref, table)
    kg = np.asarray(kgrid).reshape(C, -1)
    rel = ogrid.rel_l2(kg, gref)
    assert rel <= 1e-5, f"fp32 grid rel-L2 vs reference flow {rel:.2e}"
    dens_gpu = np.asarray(mplan.density).reshape(-1)
    # weights carry float32 exp, so density matches the oracle's to 1e-6 ...
    np.testing.assert_allclose(dens_gpu, dens, rtol=1e-6, atol=1e-7)
    # ... and is bitwise the reference's sorted-order float32 sum of the
    # plan's own sqrt weights
    p_gpu = dict(p, sqrt_weights=g.plan.sqrt_weights.cpu().numpy())
    assert np.array_equal(dens_gpu, oplan.density(p_gpu)), "density summation order differs"
    g64 = ogrid.grid_fp64(y.reshape(C, -1), p, e_ref, t64)
    assert ogrid.rel_l2(kg, g64) <= 1e-5
    # sparse output (every candidate replica)
    sp = g.interpolate(y)
    sref = ogrid.sparse_reference(y.reshape(C, -1), e_ref, table, p["offsets"].shape[0])
    np.testing.assert_allclose(np.asarray(sp), sref, rtol=1e-5, atol=1e-5 * np.abs(sref).max())
    # ret_image
    img = g.interpolate(y, ret_image=True)
    iref = ogrid.image_rss(gref, p["grid_shape"], shape)
    assert ogrid.rel_l2(np.asarray(img), iref) <= 1e-5


def test_grid_deterministic_and_batch_invariant(device):
    case = CASES[1]
    g, p, _ = _build(case, device)
    g.calc_interp_table(_calib(8, (24, 24)), lamda=1e-3, quant=32)
    rng = np.random.default_rng(3)
    y = torch.from_numpy((rng.standard_normal((8, 32, 128)) + 1j * rng.standard_normal((8, 32, 128))).astype(np.complex64))
    yd = y.to(device)
    a, _ = g.interpolate(yd, grid=True)
    b, _ = g.interpolate(yd, grid=True)
    assert torch.equal(a, b)
    y2 = torch.stack([yd * 0.5, yd])  # batched: second element equals the single run
    c, _ = g.interpolate(y2, grid=True)
    assert torch.equal(c[1], a)


@pytest.mark.parametrize("mode,tol", [("bf16", 5e-3), ("bf16x3", 1e-5)])
@pytest.mark.parametrize("ci", [1, 4])
def test_bf16_modes_vs_fp64(device, mode, tol, ci):
    case = CASES[ci]
    name, shape, mk, osf, w, ks, C, cshape, q = case
    g, p, _ = _build(case, device)
    calib = _calib(C, cshape)
    g.calc_interp_table(calib, lamda=0.01, **q)
    e_ref = g._interp_idx.reshape(-1).cpu().numpy()
    t64, _, _ = otable.table(calib, 0.01, q.get("precision", 1), q.get("quant"))
    rng = np.random.default_rng(12)
    npts = p["target_idx"].shape[:-1]
    y = (rng.standard_normal((C, *npts)) + 1j * rng.standard_normal((C, *npts))).astype(np.complex64)
    kg, _ = g.interpolate(y, grid=True, mode=mode)
    g64 = ogrid.grid_fp64(y.reshape(C, -1), p, e_ref, t64)
    rel = ogrid.rel_l2(np.asarray(kg).reshape(C, -1), g64)
    print(f"{name} {mode}: rel-L2 vs FP64 = {rel:.3e}")
    assert rel <= tol
    k2, _ = g.interpolate(y, grid=True, mode=mode)
    assert np.array_equal(np.asarray(kg), np.asarray(k2))


def test_stacked_plan_matches_per_element(device):
    coords = np.stack([_radial(12, 64, s) for s in range(3)])  # (3, 12, 64, 2) -> auto-stack
    g = GrogInterpolator((32, 32), coords, kernel_width=2, oversamp=2.0, device=device)
    assert g.plan.stack_shape == (3,)
    calib = _calib(4, (16, 16))
    g.calc_interp_table(calib, lamda=0.01)
    rng = np.random.default_rng(2)
    y = (rng.standard_normal((3, 4, 12, 64)) + 1j * rng.standard_normal((3, 4, 12, 64))).astype(np.complex64)
    ks, mp = g.interpolate(y, grid=True)
    assert ks.shape == (3, 4, 64, 64) and mp.density.shape == (3, 64, 64)
    for s in range(3):
        gs = GrogInterpolator((32, 32), coords[s], kernel_width=2, oversamp=2.0, device=device)
        gs.calc_interp_table(calib, lamda=0.01)
        np.testing.assert_array_equal(g.plan.target_idx[s].cpu().numpy(), gs.plan.target_idx.cpu().numpy())
        k1, _ = gs.interpolate(y[s], grid=True)
        assert np.array_equal(np.asarray(ks[s]), np.asarray(k1))


def test_hot_prepass_and_schedule_do_not_change_bits(device, monkeypatch):
    """The hot-cell pre-pass and the block count only redistribute work:
    the reduction order is fixed by the plan, so the grids are identical."""
    import pygrog_amd.calib._grog as grog_mod

    coords = _radial(64, 256, 9)  # DC and the first rings are hot cells
    calib = _calib(8, (24, 24))
    rng = np.random.default_rng(4)
    y = torch.from_numpy((rng.standard_normal((8, 64, 256)) + 1j * rng.standard_normal((8, 64, 256)))
                         .astype(np.complex64)).to(device)

    def build():
        g = GrogInterpolator((128, 128), coords, kernel_width=3, device=device)
        g.calc_interp_table(calib, lamda=1e-3, quant=32)
        return g

    g_hot = build()
    assert g_hot.plan._compiled.hot[0]["n_hot"] > 0
    monkeypatch.setattr(grog_mod, "HOT_MIN_UNITS", 10**9)
    g_cold = build()
    assert g_cold.plan._compiled.hot[0]["n_hot"] == 0
    yy = y.reshape(1, 1, 8, -1).contiguous()
    for mode in ("fp32", "bf16x3", "bf16"):
        a = g_hot.grid_device(yy, mode)
        b = g_cold.grid_device(yy, mode)
        assert torch.equal(a, b), mode
    # a different persistent grid size (tile -> workgroup mapping)
    comp = g_hot.plan._compiled
    out = torch.empty(1, 1, 8, comp.G, dtype=torch.complex64, device=device)
    yt = torch.empty(comp.npts * 8, dtype=torch.complex64, device=device)
    g_hot._transpose_all(yy, yt, None)
    g_hot.grid_program(yt, out, 1, n_blocks=8).launch("pg_grid_fp32")
    assert torch.equal(out, g_hot.grid_device(yy, "fp32"))


def test_plan_weights_big_cells_bit_exact(device):
    """Cells with hundreds to thousands of replicas (800 spokes through a
    16 x 16 grid: 21 cells above 256, one above 2048) take the plan's
    workgroup path for the per-cell sums (the one-call compiler lists every
    cell above 64 replicas for a kernel of one workgroup per cell; a cell
    above 2048 crosses the double-buffered chunks of its ordered density
    sum): weights and sqrt weights match the oracle and the density is
    bitwise the reference's ordered float32 sum."""
    coords = _radial(800, 32, 3)
    shape = (16, 16)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, None, 3, "circle"))
    cnt = np.bincount(p["indices"][p["sqrt_weights"] > 0], minlength=256)
    assert (cnt > 256).sum() >= 10 and cnt.max() > 2048
    np.testing.assert_array_equal(g.plan.indices.cpu().numpy(), p["indices"])
    np.testing.assert_allclose(g.plan.weights.reshape(-1).cpu().numpy(), p["weights"].reshape(-1), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(g.plan.sqrt_weights.cpu().numpy(), p["sqrt_weights"], rtol=1e-6, atol=1e-7)
    dens_gpu = g.plan._compiled.density.reshape(-1).cpu().numpy()
    p_gpu = dict(p, sqrt_weights=g.plan.sqrt_weights.cpu().numpy())
    assert np.array_equal(dens_gpu, oplan.density(p_gpu)), "density summation order differs"


def test_deferred_plan_equals_blocking_plan(device):
    """A plan compiled without waiting (pg_plan_compile_async: summary
    through page-locked memory behind an event; the serving loop fits the
    table meanwhile) holds every field the blocking call produces, bit for
    bit, on a 2D radial and a 3D kooshball trajectory."""
    from pygrog_amd import synth

    cases = [((96, 96), _radial(48, 160, seed=3), None, 3, "circle"),
             ((32, 32, 32), synth.kooshball_3d(300, 48), (1, 1, 1), 2, "sphere")]
    for shape, coords, over, kw, ks in cases:
        a = GrogInterpolator(shape, coords, over, kw, ks, device=device)
        pending = GrogInterpolator._launch_plan(shape, coords, over, kw, ks, device=device, defer=True)
        b = GrogInterpolator._from_pending(pending)
        ca, cb = a.plan._compiled, b.plan._compiled
        assert (ca.n_inv, ca.n_valid, ca.G, ca.npts) == (cb.n_inv, cb.n_valid, cb.G, cb.npts)
        for name in ("target_idx", "weights", "indices", "sqrt_weights", "sort_perm", "inv_perm"):
            assert torch.equal(getattr(a.plan, name), getattr(b.plan, name)), name
        for name in ("cell_ptr", "cell_unit", "unit_cell", "unit_rep", "unit_hot", "cell_vrow", "density",
                     "n_invalid", "distances"):
            assert torch.equal(getattr(ca, name), getattr(cb, name)), name
        R = ca.n_valid[0]  # the record arrays are sized by the candidates; the valid replicas come first
        for name in ("rep_src", "rep_s"):
            assert torch.equal(getattr(ca, name)[:, :R], getattr(cb, name)[:, :R]), name
        assert [(h["n_hot"], h["n_hot_units"]) for h in ca.hot] == [(h["n_hot"], h["n_hot_units"]) for h in cb.hot]
