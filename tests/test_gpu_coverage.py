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
    assert ogrid.rel_l2(g.grid_device(yy, "bf16x3c").reshape(C, -1).cpu().numpy(), g64) <= 2e-4
    assert ogrid.rel_l2(g.grid_device(yy, "bf16").reshape(C, -1).cpu().numpy(), g64) <= 5e-3


@pytest.mark.parametrize("C", [1, 3, 16, 17, 32, 33, 64])
def test_coil_counts_every_kernel_shape(device, C):
    """Every coil group (4, 8, 16, 32, 64 lanes) and MFMA shape, odd and
    even coil counts (unaligned rows), against the oracle; 16 and 32 coils
    are the grouped apply's complex64-row (MT=2, KS=1) and 24-bit-row
    (MT=4, KS=2) instances of the benchmark workloads."""
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
    assert ogrid.rel_l2(g.grid_device(yy, "bf16x3c").reshape(C, -1).cpu().numpy(), g64) <= 2e-4
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


@pytest.mark.parametrize("C", [3, 16, 18, 32, 33, 64])
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


@pytest.mark.parametrize("C", [8, 16, 32])
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


@pytest.mark.parametrize("C", [18, 32])
def test_reduce_unaligned_v_same_bits(device, C):
    """A caller's V buffer need only be 8-byte aligned: the grouped path's
    V moved to an 8-byte offset reduces to the same bits (16-byte row loads
    at 8-byte alignment), with the hot pre-pass, on 256-cell wave tiles."""
    from pygrog_amd.calib import GrogInterpolator
    from pygrog_amd.calib._grog import GridProgram

    rng = np.random.default_rng(300 + C)
    shape = (80, 80)
    coords = _radial_full(40, 160, seed=C)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    calib = (rng.standard_normal((C, 16, 16)) + 1j * rng.standard_normal((C, 16, 16))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=1e-2, quant=16)
    assert g.plan._compiled.hot[0]["n_hot"] > 0
    n = coords.shape[0] * coords.shape[1]
    y = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    yy = torch.from_numpy(y).to(device).reshape(1, 1, C, -1).contiguous()
    g._tuned = {"bf16x3": (256, None)}
    cache = {}
    base = g.grid_device(yy, "bf16x3", program_cache=cache).clone()
    (vbuf, yt, _), = cache.values()
    vf = vbuf.view(torch.float32).reshape(-1)
    moved = torch.empty(vf.numel() + 4, dtype=torch.float32, device=device)
    assert moved.data_ptr() % 16 == 0
    moved[2:2 + vf.numel()] = vf

    class _At8:
        def data_ptr(self):
            return moved.data_ptr() + 8

    out = torch.full_like(base, complex(float("nan"), 0.0))
    comp = g.plan._compiled
    prog = GridProgram(g._frames(yt, out, 1, _At8()), comp.C, comp.G, comp.device, 256, None)
    prog.launch("pg_grid_reduce_v")
    torch.cuda.synchronize()
    assert torch.equal(out, base)


def _s24_expected(x: np.ndarray) -> np.ndarray:
    """float32 -> the 24-bit layout's value as float32: nearest-even multiple
    of 2^(e - 15) (e: the binade exponent, -126 for subnormals), computed in
    float64; NaN stays NaN."""
    x64 = x.astype(np.float64)
    fin = np.isfinite(x64) & (x64 != 0)
    e = np.full(x64.shape, -126.0)
    e[fin] = np.maximum(np.floor(np.log2(np.abs(x64[fin]))), -126.0)
    q = np.exp2(e - 15.0)
    out = np.where(fin, np.rint(x64 / q) * q, x64)
    with np.errstate(over="ignore"):
        return out.astype(np.float32)


def _s24_decode(buf: np.ndarray, N: int, C: int, rb: int) -> np.ndarray:
    b = buf.reshape(N, rb)[:, :6 * C].reshape(N, 2 * C, 3).astype(np.uint32)
    return ((b[..., 0] << 8) | (b[..., 1] << 16) | (b[..., 2] << 24)).view(np.float32)


@pytest.mark.parametrize("C,N", [(32, 4096), (16, 1000), (3, 77), (40, 130)])
def test_transpose_s24_rows(device, C, N):
    """pg_transpose_cn_s24 (full 64 x 32 tiles and ragged ones): every value
    of the (N, C) sample-major rows is the fp32 input rounded to nearest even
    at 16 significant bits, stored little-endian in 3 bytes, rows of
    pg_s24_row_bytes(C) bytes -- zeros, infinities, NaN, subnormals, ties and
    a round-up into infinity included; pg_transpose_frames_s24 writes each
    frame's rows at row_bytes x its sample offset."""
    from pygrog_amd import _lib

    lib = _lib.load()
    rb = int(lib.pg_s24_row_bytes(C))
    assert rb == 16 * ((6 * C + 15) // 16)
    rng = np.random.default_rng(C * 7 + N)
    re = (rng.standard_normal((C, N)) * 10.0 ** rng.integers(-30, 30, (C, N))).astype(np.float32)
    im = rng.standard_normal((C, N)).astype(np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-40, -3e-39, 3.4028235e38,
                        np.float32(1 + 2 ** -16), np.float32(1 + 3 * 2 ** -16), np.float32(1 + 2 ** -16 + 2 ** -23),
                        np.float32(2 - 2 ** -17)], dtype=np.float32)
    re.reshape(-1)[:special.size] = special
    im.reshape(-1)[-special.size:] = special
    y = np.empty((C, N), dtype=np.complex64)  # parts set apart: 1j * inf would leak NaN into the real part
    y.real, y.imag = re, im
    yd = torch.from_numpy(y).to(device)
    ys = torch.zeros(N * rb, dtype=torch.uint8, device=device)
    _lib.call("pg_transpose_cn_s24", _lib.ptr(yd), C, N, _lib.ptr(ys), _lib.stream(device))
    got = _s24_decode(ys.cpu().numpy(), N, C, rb)
    want = _s24_expected(np.stack([re.T, im.T], axis=-1).reshape(N, 2 * C))
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), want[~nan].view(np.uint32))
    # two frames back to back
    n0 = N // 3
    off = torch.tensor([0, n0, N], dtype=torch.int64, device=device)
    yf = torch.from_numpy(np.concatenate([y[:, :n0].reshape(-1), y[:, n0:].reshape(-1)])).to(device)
    ysf = torch.zeros(N * rb, dtype=torch.uint8, device=device)
    _lib.call("pg_transpose_frames_s24", _lib.ptr(yf), _lib.ptr(off), 2, N - n0, C, _lib.ptr(ysf), _lib.stream(device))
    gf = _s24_decode(ysf.cpu().numpy(), N, C, rb)
    assert np.array_equal(np.isnan(gf), nan)
    assert np.array_equal(gf[~nan].view(np.uint32), want[~nan].view(np.uint32))


@pytest.mark.parametrize("C", [20, 32])
def test_grouped_s24_same_bits_on_16bit_inputs(device, monkeypatch, C):
    """For k-space values that already have 16 significant bits the 24-bit
    sample rows hold them exactly and the split x = hi + lo is exact on both
    paths, so the BF16 modes grid the same bits from 24-bit rows as from
    complex64 rows."""
    import pygrog_amd.calib._grog as grog_mod
    from pygrog_amd.calib import GrogInterpolator

    rng = np.random.default_rng(500 + C)
    shape = (64, 64)
    coords = _radial_full(48, 128, seed=C)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    calib = (rng.standard_normal((C, 16, 16)) + 1j * rng.standard_normal((C, 16, 16))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=1e-2, quant=16)
    n = coords.shape[0] * coords.shape[1]
    y = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    y16 = (_s24_expected(y.real) + 1j * _s24_expected(y.imag)).astype(np.complex64)
    yy = torch.from_numpy(y16).to(device).reshape(1, 1, C, -1).contiguous()
    assert grog_mod.uses_s24("bf16x3", C)
    for mode in ("bf16x3", "bf16"):
        a = g.grid_device(yy, mode, program_cache={}).clone()
        with monkeypatch.context() as m:
            m.setattr(grog_mod, "S24_MAX_COILS", 0)
            b = g.grid_device(yy, mode, program_cache={})
        assert torch.equal(a, b), mode


@pytest.mark.parametrize("C", [8, 20])
def test_grid_device_batch_of_stack_bf16(device, C):
    """grid_device over a batch of 2 of a stacked plan of 3 elements, BF16
    modes (20 coils: 24-bit sample rows, 8: complex64 rows; bf16x3c:
    block-floating-point sample and V rows, whose frames sit at their own
    row size): every (batch, element) grid equals that element gridded
    alone, bit for bit."""
    from pygrog_amd.calib import GrogInterpolator

    rng = np.random.default_rng(700 + C)
    shape = (40, 40)
    coords = np.stack([_radial_full(6, 80, seed=10 * C + i) for i in range(3)])
    calib = (rng.standard_normal((C, 16, 16)) + 1j * rng.standard_normal((C, 16, 16))).astype(np.complex64)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    assert g.plan.stack_shape == (3,)
    g.calc_interp_table(calib, lamda=1e-2, quant=16)
    n = coords.shape[1] * coords.shape[2]
    y = (rng.standard_normal((2, 3, C, n)) + 1j * rng.standard_normal((2, 3, C, n))).astype(np.complex64)
    yy = torch.from_numpy(y).to(device).contiguous()
    for mode in ("bf16x3", "bf16", "bf16x3c"):
        got = g.grid_device(yy, mode, program_cache={})
        for s in range(3):
            gs = GrogInterpolator(shape, coords[s], kernel_width=3, device=device)
            gs.calc_interp_table(calib, lamda=1e-2, quant=16)
            for b in range(2):
                alone = gs.grid_device(yy[b, s].reshape(1, 1, C, -1).contiguous(), mode)
                assert torch.equal(got[b, s], alone[0, 0]), (mode, b, s)


@pytest.mark.parametrize("case", ["2d", "3d", "stack", "nohot", "allhot"])
def test_plan_hot_rows_match_torch_lists(device, monkeypatch, case):
    """The native plan compiler's reduction units, hot lists and row
    numbering (pg_plan_compile) equal the torch formulation
    (compile_units, hot_lists, ordinary_rows) entry for entry: 2D with a hot
    DC cell, a 3D kooshball, a stack of 3 elements, no hot cells, every
    occupied cell hot."""
    import pygrog_amd.calib._grog as G
    from pygrog_amd import synth
    from pygrog_amd.calib import GrogInterpolator

    mu = {"nohot": 10**9, "allhot": 1}.get(case, G.HOT_MIN_UNITS)
    monkeypatch.setattr(G, "HOT_MIN_UNITS", mu)
    if case == "3d":
        g = GrogInterpolator((24, 24, 24), synth.kooshball_3d(300, 32), oversamp=(1, 1, 1), kernel_width=2,
                             kernel_shape="sphere", device=device)
    elif case == "stack":
        g = GrogInterpolator((40, 40), np.stack([_radial_full(40, 96, seed=s) for s in range(3)]), kernel_width=3,
                             device=device)
        assert g.plan._compiled.S == 3
    else:
        g = GrogInterpolator((80, 80), _radial_full(40, 160, seed=1), kernel_width=3, device=device)
    comp = g.plan._compiled
    S, Gc = comp.S, comp.G
    cu, uc, ur = G.compile_units(comp.cell_ptr, S, Gc, device, G.UNIT_CHUNK)
    assert torch.equal(comp.cell_unit, cu) and torch.equal(comp.unit_rep, ur)
    assert torch.equal(comp.unit_cell[:uc.numel()], uc) or uc.numel() == 1
    uh2, per2 = G.hot_lists(comp.cell_unit, S, Gc, device, mu)
    vrow2 = G.ordinary_rows(comp.cell_ptr, comp.cell_unit, S, Gc, mu)
    assert torch.equal(comp.cell_vrow, vrow2)
    assert torch.equal(comp.unit_hot, uh2)
    total_hot = 0
    for a, b in zip(comp.hot, per2):
        assert (a["n_hot"], a["n_hot_units"]) == (b["n_hot"], b["n_hot_units"])
        assert torch.equal(a["hot_ptr"], b["hot_ptr"])
        if a["n_hot_units"]:
            assert torch.equal(a["hot_units"], b["hot_units"])
        total_hot += a["n_hot"]
    assert (total_hot == 0) == (case == "nohot")


@pytest.mark.parametrize("R,K,chunk", [(1, 1, 128), (1000, 7, 128), (250000, 1089, 128), (100000, 68921, 128),
                                       (5000, 3, 1), (70000, 2, 128)])
def test_group_by_entry_matches_torch(device, R, K, chunk):
    """pg_group_by_entry (one call) against the torch formulation of the same
    contract: replicas in a stable sort by entry, each entry's range, the
    work list of (entry, offset) chunks in entry order, the gathered sources
    and weights -- equal entry for entry; skewed entry counts included."""
    from pygrog_amd.calib._grog import group_by_entry

    g = torch.Generator(device="cpu").manual_seed(R + K)
    # a few entries carry most replicas, as the zero-shift entries do
    heavy = torch.randint(0, K, (max(K // 50, 1),), generator=g)
    ent = torch.where(torch.rand(R, generator=g) < 0.5, heavy[torch.randint(0, heavy.numel(), (R,), generator=g)],
                      torch.randint(0, K, (R,), generator=g)).to(torch.int32)
    src = torch.randint(0, 10**6, (R,), generator=g, dtype=torch.int32)
    sw = torch.rand(R, generator=g)
    grp = group_by_entry(ent.to(device), R, K, device, chunk, src.to(device), sw.to(device))
    order = torch.sort(ent.to(torch.int64), stable=True).indices
    counts = torch.bincount(ent.to(torch.int64), minlength=K)
    ptr = torch.cat([torch.zeros(1, dtype=torch.int64), torch.cumsum(counts, 0)])
    chunks = (counts + chunk - 1) // chunk
    total = int(chunks.sum())
    e = torch.repeat_interleave(torch.arange(K), chunks)
    first = torch.cumsum(chunks, 0) - chunks
    off = (torch.arange(total) - first[e]) * chunk
    assert grp.n_items == total
    assert torch.equal(grp.ent_ptr.cpu().to(torch.int64), ptr)
    assert torch.equal(grp.ent_rep.cpu().to(torch.int64), order)
    assert torch.equal(grp.items[:total].cpu().to(torch.int64), torch.stack([e, off], 1))
    assert torch.equal(grp.ent_src.cpu(), src[order])
    assert torch.equal(grp.ent_s.cpu(), sw[order])


@pytest.mark.parametrize("R,K", [(1, 1), (250000, 1089), (100000, 68921)])
def test_group_by_entry_deferred(device, R, K):
    """pg_group_by_entry_async: the same groups as the synchronising call,
    the work list's unused capacity marked with entry 0xFFFFFFFF, and a
    grouped apply launched over the whole capacity writing the same V bits
    as one launched over the exact length."""
    from pygrog_amd import _lib
    from pygrog_amd.calib._grog import group_by_entry

    g = torch.Generator(device="cpu").manual_seed(7 * R + K)
    ent = torch.randint(0, K, (R,), generator=g, dtype=torch.int32)
    src = torch.randint(0, 4096, (R,), generator=g, dtype=torch.int32)
    sw = torch.rand(R, generator=g)
    a = group_by_entry(ent.to(device), R, K, device, 128, src.to(device), sw.to(device))
    b = group_by_entry(ent.to(device), R, K, device, 128, src.to(device), sw.to(device), defer=True)
    torch.cuda.synchronize()
    assert b.n_items == b.items.shape[0] >= a.n_items
    for x, y in ((a.ent_ptr, b.ent_ptr), (a.ent_rep, b.ent_rep), (a.ent_src, b.ent_src), (a.ent_s, b.ent_s)):
        assert torch.equal(x, y)
    assert torch.equal(a.items[:a.n_items], b.items[:a.n_items])
    assert bool((b.items[a.n_items:, 0] == -1).all()) and bool((b.items[a.n_items:, 1] == 0).all())
    # the apply over the capacity: same V as over the exact length
    C = 8
    n_src = 4096
    y = (torch.randn(C, n_src, generator=g) + 1j * torch.randn(C, n_src, generator=g)).to(torch.complex64)
    yt = y.T.contiguous().to(device)
    tbl = (torch.randn(K, C, C, generator=g) + 1j * torch.randn(K, C, C, generator=g)).to(torch.complex64).to(device)
    n = int(_lib.load().pg_table_bf16_elems(K, C, 2))
    packed = torch.empty(n, dtype=torch.int16, device=device)
    _lib.call("pg_table_pack_bf16", _lib.ptr(tbl), K, C, 2, _lib.ptr(packed), _lib.stream(device))
    outs = []
    for grp in (a, b):
        v = torch.full((R * C,), complex(float("nan"), 0.0), dtype=torch.complex64, device=device)
        _lib.call("pg_apply_grouped_bf16", _lib.ptr(yt), _lib.ptr(packed), _lib.ptr(grp.ent_ptr), _lib.ptr(grp.ent_rep),
                  _lib.ptr(grp.ent_src), _lib.ptr(grp.ent_s), _lib.ptr(grp.items), grp.n_items, K, C, 3,
                  _lib.ptr(v), _lib.stream(device))
        outs.append(v)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    assert not bool(torch.isnan(outs[0].real).any())


@pytest.mark.parametrize("ndim,n,C,order", [(2, 33, 8, "reference"), (3, 11, 16, "reference"),
                                            (3, 5, 16, "corrected"), (3, 3, 32, "reference"), (2, 4, 1, "reference")])
def test_table_compose_matches_batched_products(device, ndim, n, C, order):
    """pg_table_compose_f64 against the batched complex128 products of
    prepare_grappa_table (torch on the device): the FP64 tables within a few
    units in the last place (different summation order), the complex64
    tables their rounding."""
    from pygrog_amd.calib._grappa import compose_table, prepare_grappa_table

    g = torch.Generator(device="cpu").manual_seed(ndim * 100 + n + C)
    P = (torch.randn(ndim, n, C, C, generator=g, dtype=torch.float64)
         + 1j * torch.randn(ndim, n, C, C, generator=g, dtype=torch.float64)).to(device)
    t32, t64 = compose_table(P, ndim, n, order, keep_fp64=True)
    ref = prepare_grappa_table(P[0], P[1], P[2] if ndim == 3 else None, n, ndim, axis_order=order)
    torch.cuda.synchronize()
    assert t64.shape == ref.shape == (n ** ndim, C, C)
    err = (t64 - ref).abs().max().item() / ref.abs().max().item()
    assert err < 1e-14, err
    assert torch.allclose(t32, t64.to(torch.complex64), rtol=0, atol=0)
    assert (t32 != ref.to(torch.complex64)).float().mean().item() < 1e-3
