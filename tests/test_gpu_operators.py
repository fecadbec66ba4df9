"""SparseFFT / MaskedFFT / GrogToeplitzOp / op.solve / GrogInterpolator on
the HIP device, following the reference's own operator tests
(tests/test_sparse_fft.py, test_masked_fft.py, test_toeplitz.py:128-150,
test_solve.py, test_grog_interpolator.py) and checked against the oracle's
float64 restatements (oracle/operators.py) at rel-L2 <= 1e-5 (FP32 path).
"""

import numpy as np
import pytest
import torch

from oracle import grid as ogrid
from oracle import operators as oop

pytestmark = pytest.mark.gpu


def _rel(a, b) -> float:
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _cdot(a, b):
    return torch.vdot(a.flatten(), b.flatten())


def _np(t):
    return t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def _make_op(grid_shape, image_shape, n, *, smaps=None, seed=0, weights=None, toeplitz=None):
    from pygrog_amd.operator import SparseFFT

    rng = np.random.default_rng(seed)
    G = int(np.prod(grid_shape))
    idx = rng.integers(0, G, size=n)
    w = np.ones(n, np.float32) if weights is None else weights
    return SparseFFT(grid_shape, image_shape, idx, w, smaps=smaps, toeplitz=toeplitz), idx, w


def _perm(idx):
    p = np.argsort(idx, kind="stable")
    inv = np.empty_like(p)
    inv[p] = np.arange(p.size)
    return p, inv


# ---------------------------------------------------------------------------
# SparseFFT
# ---------------------------------------------------------------------------
def test_sparse_fft_shapes_devices_and_call(device):
    op, _, _ = _make_op((32, 32), (24, 24), 256)
    ksp = torch.randn(4, 256, dtype=torch.complex64, device=device)
    img = torch.randn(4, 24, 24, dtype=torch.complex64, device=device)
    a = op.adjoint(ksp)
    f = op.forward(img)
    assert a.shape == (4, 24, 24) and a.device.type == "cuda"
    assert f.shape == (4, 256) and f.device.type == "cuda"
    torch.testing.assert_close(op(img), f)
    torch.testing.assert_close(op(ksp, adjoint=True), a)
    # host tensors in -> host tensors out, computed on the device
    assert op.adjoint(ksp.cpu()).device.type == "cpu"
    smaps = torch.randn(4, 24, 24, dtype=torch.complex64)
    op_s, _, _ = _make_op((32, 32), (24, 24), 256, smaps=smaps)
    assert op_s.adjoint(ksp).shape == (24, 24)
    assert op_s.forward(torch.randn(24, 24, dtype=torch.complex64, device=device)).shape == (4, 256)
    op3, _, _ = _make_op((16, 16, 16), (12, 12, 12), 128)
    assert op3.adjoint(torch.randn(2, 128, dtype=torch.complex64, device=device)).shape == (2, 12, 12, 12)


@pytest.mark.parametrize("grid,image,n,with_smaps,weighted", [
    ((32, 32), (24, 24), 256, False, False),
    ((32, 32), (24, 24), 2000, True, True),     # many collisions per cell
    ((16, 16, 16), (12, 12, 12), 700, False, True),
    ((40, 36), (40, 36), 999, True, False),     # no oversampling, ragged n
])
def test_sparse_fft_vs_oracle(device, grid, image, n, with_smaps, weighted):
    rng = np.random.default_rng(n)
    C = 3
    smaps = (rng.standard_normal((C, *image)) + 1j * rng.standard_normal((C, *image))).astype(np.complex64) \
        if with_smaps else None
    w = rng.uniform(0.5, 2.0, n).astype(np.float32) if weighted else None
    op, idx, w = _make_op(grid, image, n, smaps=None if smaps is None else torch.as_tensor(smaps), seed=n, weights=w)
    p, inv = _perm(idx)
    assert np.array_equal(_np(op.sort_perm), p)
    sqw = np.sqrt(w.astype(np.float32))[p]
    x = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    got = _np(op.adjoint(torch.as_tensor(x).to(device)))
    ref = oop.sparse_adjoint(x, idx[p], sqw, p, grid, image, smaps)
    assert _rel(got, ref) <= 1e-5
    im = (rng.standard_normal((C, *image)) + 1j * rng.standard_normal((C, *image))).astype(np.complex64)
    if smaps is not None:
        im = im[0]
    got_f = _np(op.forward(torch.as_tensor(im).to(device)))
    ref_f = oop.sparse_forward(im, idx[p], sqw, inv, grid, image, smaps)
    assert _rel(got_f, ref_f) <= 1e-5


@pytest.mark.parametrize("grid,image,n,smaps", [((32, 32), (24, 24), 256, False), ((32, 32), (24, 24), 256, True),
                                                ((16, 16, 16), (12, 12, 12), 128, False)])
def test_sparse_fft_adjointness(device, grid, image, n, smaps):
    torch.manual_seed(0)
    rng = np.random.default_rng(42)
    sm = torch.randn(4, *image, dtype=torch.complex64) if smaps else None
    op, _, _ = _make_op(grid, image, n, smaps=sm, seed=3, weights=rng.uniform(0.5, 2.0, n).astype(np.float32))
    x = torch.randn(4, n, dtype=torch.complex64, device=device)
    y = torch.randn(*image, dtype=torch.complex64, device=device) if smaps else \
        torch.randn(4, *image, dtype=torch.complex64, device=device)
    torch.testing.assert_close(_cdot(op.adjoint(x), y), _cdot(x, op.forward(y)), rtol=1e-4, atol=1e-5)


def test_sparse_fft_out_of_range_indices_are_skipped(device):
    """Samples outside [0, G) do not scatter and gather as 0 (reference
    sparse_ops_cpu_impl.inl:98, 161)."""
    from pygrog_amd.operator import SparseFFT

    rng = np.random.default_rng(5)
    G = 16 * 16
    idx = rng.integers(0, G, size=300)
    idx[::7] = G + 3
    idx[3::11] = -1
    w = np.ones(300, np.float32)
    op = SparseFFT((16, 16), (12, 12), idx, w)
    x = (rng.standard_normal((2, 300)) + 1j * rng.standard_normal((2, 300))).astype(np.complex64)
    keep = (idx >= 0) & (idx < G)
    op2 = SparseFFT((16, 16), (12, 12), idx[keep], w[keep])
    a = _np(op.adjoint(torch.as_tensor(x).to(device)))
    b = _np(op2.adjoint(torch.as_tensor(x[:, keep]).to(device)))
    assert _rel(a, b) <= 1e-6
    im = torch.randn(2, 12, 12, dtype=torch.complex64, device=device)
    f = _np(op.forward(im))
    assert np.all(f[:, ~keep] == 0)
    assert _rel(f[:, keep], _np(op2.forward(im))) <= 1e-6


def test_batch_helpers_vs_loop_oracle_and_adjoint(device):
    from pygrog_amd._base._fftc import fft, ifft
    from pygrog_amd.operator import gather, scatter_add

    op, idx, w = _make_op((32, 32), (24, 24), 256, seed=4)
    ksp = torch.randn(6, 256, dtype=torch.complex64, device=device)
    out = op._scatter_ifft_crop_batch(ksp)
    sp, ix, sw = op.sort_perm.to(device), op.indices.to(device), op.sqrt_weights.to(device)
    loop = []
    for b in range(6):
        g = torch.zeros(op.grid_size, dtype=torch.complex64, device=device)
        scatter_add(g, ksp[b][sp], ix, sw)
        loop.append(ifft(g.reshape(op.grid_shape), axes=op.fft_axes)[op._pad_slices])
    torch.testing.assert_close(out, torch.stack(loop), rtol=1e-5, atol=1e-5)
    p, _ = _perm(idx)
    assert _rel(_np(out), oop.sparse_adjoint(_np(ksp), idx[p], np.sqrt(w)[p], p, (32, 32), (24, 24))) <= 1e-5

    imgs = torch.randn(6, 24, 24, dtype=torch.complex64, device=device)
    outf = op._fft_pad_gather_batch(imgs)
    inv = op.inv_perm.to(device)
    loopf = []
    for b in range(6):
        pad = torch.zeros(*op.grid_shape, dtype=torch.complex64, device=device)
        pad[op._pad_slices] = imgs[b]
        loopf.append(gather(fft(pad, axes=op.fft_axes).reshape(-1), ix, sw)[inv])
    torch.testing.assert_close(outf, torch.stack(loopf), rtol=1e-5, atol=1e-5)

    x = torch.randn(5, 256, dtype=torch.complex64, device=device)
    y = torch.randn(5, 24, 24, dtype=torch.complex64, device=device)
    torch.testing.assert_close(_cdot(op._scatter_ifft_crop_batch(x), y), _cdot(x, op._fft_pad_gather_batch(y)),
                               rtol=1e-4, atol=1e-5)


def test_sparse_fft_many_coils_chunked(device):
    """More than 64 rows go through the fused scatter in coil chunks."""
    op, idx, w = _make_op((24, 24), (20, 20), 500, seed=9)
    rng = np.random.default_rng(9)
    x = (rng.standard_normal((70, 500)) + 1j * rng.standard_normal((70, 500))).astype(np.complex64)
    p, _ = _perm(idx)
    got = _np(op.adjoint(torch.as_tensor(x).to(device)))
    assert _rel(got, oop.sparse_adjoint(x, idx[p], np.sqrt(w)[p], p, (24, 24), (20, 20))) <= 1e-5


# ---------------------------------------------------------------------------
# GROG set-up shared by the MaskedFFT / solve / interpolator tests
# (reference tests/test_masked_fft.py:15-53)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def grog_setup():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from pygrog_amd.calib import GrogInterpolator

    rng = np.random.default_rng(42)
    shape, image_shape = (16, 16), (12, 12)
    nc, nv, nr = 3, 8, 16
    coords = rng.standard_normal((nv, nr, 2)).astype(np.float32)
    g = GrogInterpolator(shape=shape, coords=coords, kernel_width=2, oversamp=2.0, image_shape=image_shape)
    calib = (rng.standard_normal((nc, *shape)) + 1j * rng.standard_normal((nc, *shape))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=0.01, precision=1)
    data = (rng.standard_normal((nc, nv, nr)) + 1j * rng.standard_normal((nc, nv, nr))).astype(np.complex64)
    smaps = (rng.standard_normal((nc, *image_shape)) + 1j * rng.standard_normal((nc, *image_shape))).astype(
        np.complex64)
    return {"grog": g, "data": data, "smaps": torch.as_tensor(smaps), "image_shape": image_shape, "n_coils": nc}


def test_interpolate_grid_returns_plan_and_kspace(grog_setup):
    from pygrog_amd.operator import MaskedFFTPlan

    g = grog_setup["grog"]
    kgrid, plan = g.interpolate(grog_setup["data"], grid=True)
    assert isinstance(kgrid, np.ndarray) and isinstance(plan, MaskedFFTPlan)
    assert kgrid.shape == (grog_setup["n_coils"], *plan.grid_shape)
    assert np.array_equal(np.asarray(plan.mask), (np.asarray(plan.density) > 0).astype(np.float32))
    with pytest.raises(ValueError):
        g.interpolate(grog_setup["data"], grid=True, ret_image=True)


def test_maskedfft_shapes_and_plan_constructor(grog_setup):
    from pygrog_amd.operator import MaskedFFT

    g = grog_setup["grog"]
    kgrid, plan = g.interpolate(grog_setup["data"], grid=True)
    op = MaskedFFT(plan=plan, smaps=grog_setup["smaps"])
    op2 = MaskedFFT(grid_shape=plan.grid_shape, image_shape=plan.image_shape, mask=plan.mask, density=plan.density,
                    smaps=grog_setup["smaps"])
    kt = torch.as_tensor(kgrid)
    torch.testing.assert_close(op.adjoint(kt), op2.adjoint(kt))
    x = torch.randn(*grog_setup["image_shape"], dtype=torch.complex64)
    assert op.forward(x).shape == (grog_setup["n_coils"], *plan.grid_shape)
    opn = MaskedFFT(plan=plan)
    assert opn.forward(torch.randn(3, *grog_setup["image_shape"], dtype=torch.complex64)).shape == kt.shape
    assert opn.adjoint(kt).shape == (3, *grog_setup["image_shape"])


def test_maskedfft_vs_oracle_and_dot_product(grog_setup):
    from pygrog_amd.operator import MaskedFFT

    g = grog_setup["grog"]
    kgrid, plan = g.interpolate(grog_setup["data"], grid=True)
    sm = grog_setup["smaps"]
    op = MaskedFFT(plan=plan, smaps=sm)
    mask = np.asarray(plan.mask)
    got = _np(op.adjoint(torch.as_tensor(kgrid)))
    assert _rel(got, oop.masked_adjoint(kgrid, mask, grog_setup["image_shape"], _np(sm))) <= 1e-5
    torch.manual_seed(0)
    x = torch.randn(*grog_setup["image_shape"], dtype=torch.complex64)
    y = torch.randn(grog_setup["n_coils"], *plan.grid_shape, dtype=torch.complex64)
    Ax = op.forward(x)
    assert _rel(_np(Ax), oop.masked_forward(_np(x), mask, grog_setup["image_shape"], _np(sm))) <= 1e-5
    assert torch.allclose(_cdot(Ax, y), _cdot(x, op.adjoint(y)), rtol=1e-5, atol=1e-5)


def test_maskedfft_toeplitz_matches_compose_and_batched(grog_setup):
    from pygrog_amd.operator import MaskedFFT

    g = grog_setup["grog"]
    kgrid, plan = g.interpolate(grog_setup["data"], grid=True)
    op = MaskedFFT(plan=plan, smaps=grog_setup["smaps"])
    assert op.toeplitz is True
    img = torch.randn(*grog_setup["image_shape"], dtype=torch.complex64)
    yt = op.normal(img)
    op.toeplitz = False
    op._toep_op = None
    yc = op.normal(img)
    assert torch.allclose(yt, yc, rtol=1e-4, atol=1e-4)
    kt = torch.as_tensor(kgrid)
    b = op.adjoint(torch.stack([kt, kt + 0.1]))
    assert b.shape == (2, *grog_setup["image_shape"])
    torch.testing.assert_close(b[0], op.adjoint(kt), rtol=1e-5, atol=1e-6)


def test_sparsefft_vs_maskedfft_image_close(grog_setup):
    from pygrog_amd.operator import MaskedFFT, SparseFFT

    g, data, sm = grog_setup["grog"], grog_setup["data"], grog_setup["smaps"]
    sparse_t = torch.as_tensor(np.asarray(g.interpolate(data, ret_image=False)))
    op_sp = SparseFFT(plan=g.plan, smaps=sm)
    img_sp = op_sp.adjoint(sparse_t * torch.as_tensor(g.plan.pre_weights))
    kgrid, plan = g.interpolate(data, grid=True)
    img_m = MaskedFFT(plan=plan, smaps=sm).adjoint(kgrid)
    img_m = torch.as_tensor(img_m)
    assert (torch.linalg.vector_norm(img_sp - img_m) / torch.linalg.vector_norm(img_sp)).item() < 1e-4


# ---------------------------------------------------------------------------
# GrogToeplitzOp (reference tests/test_toeplitz.py:41-150)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("osf", [1, 2])
@pytest.mark.parametrize("with_smaps", [False, True])
def test_sparse_fft_normal_matches_nested_and_oracle(device, osf, with_smaps):
    from pygrog_amd.operator import SparseFFT

    grid = (32 * osf, 32 * osf)
    image = tuple(s // osf for s in grid)
    rng = np.random.default_rng(0)
    G = int(np.prod(grid))
    idx = rng.integers(0, G, size=600)
    w = (rng.random(600).astype(np.float32) + 0.1)
    sm = torch.from_numpy(((rng.standard_normal((4, *image)) + 1j * rng.standard_normal((4, *image))) * 0.5)
                          .astype(np.complex64)) if with_smaps else None
    op_t = SparseFFT(grid, image, idx, w, smaps=sm, toeplitz=True)
    op_n = SparseFFT(grid, image, idx, w, smaps=sm, toeplitz=False)
    x = torch.randn(*image, dtype=torch.complex64) if with_smaps else torch.randn(1, *image, dtype=torch.complex64)
    yt, yn = op_t.normal(x), op_n.normal(x)
    assert (yt - yn).abs().max().item() / (yn.abs().max().item() + 1e-30) < 1e-4
    p, inv = _perm(idx)
    sqw = np.sqrt(w)[p]
    ref = oop.sparse_adjoint(oop.sparse_forward(_np(x), idx[p], sqw, inv, grid, image, None if sm is None else _np(sm)),
                             idx[p], sqw, p, grid, image, None if sm is None else _np(sm))
    assert _rel(_np(yt), ref) <= 1e-5
    assert _rel(_np(yn), ref) <= 1e-5


def test_toeplitz_self_adjoint_psd_and_3d(device):
    from pygrog_amd._toep import GrogToeplitzOp

    op, _, _ = _make_op((32, 32), (16, 16), 400, seed=11, toeplitz=True)
    T = GrogToeplitzOp(op)
    torch.manual_seed(0)
    x = torch.randn(1, 16, 16, dtype=torch.complex64, device=device)
    y = torch.randn(1, 16, 16, dtype=torch.complex64, device=device)
    assert T(x).shape == x.shape and T(x).device.type == "cuda"
    torch.testing.assert_close(_cdot(T(x), y), _cdot(x, T(y)), rtol=1e-4, atol=1e-5)
    assert _cdot(T(x), x).real.item() >= -1e-4
    op3, _, _ = _make_op((16, 16, 16), (8, 8, 8), 300, seed=12, toeplitz=True)
    x3 = torch.randn(2, 8, 8, 8, dtype=torch.complex64, device=device)
    assert op3.normal(x3).shape == x3.shape
    # the PSF is the per-cell sum of squared sqrt-weights
    ix = op.indices.to(device)
    ref = torch.zeros(op.grid_size, dtype=torch.float32, device=device)
    ref.index_add_(0, ix, op.sqrt_weights.to(device) ** 2)
    torch.testing.assert_close(T.psf.reshape(-1), ref, rtol=1e-6, atol=1e-7)


def test_sparse_fft_toeplitz_auto_policy():
    from pygrog_amd.operator import SparseFFT

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    idx = np.arange(100) % 64
    assert SparseFFT((8, 8), (8, 8), idx, np.ones(100, np.float32), toeplitz=None).toeplitz is True
    assert SparseFFT((8, 8), (8, 8), idx, np.ones(100, np.float32), device="cuda",
                     toeplitz=None).toeplitz is False


# ---------------------------------------------------------------------------
# op.solve (reference tests/test_solve.py)
# ---------------------------------------------------------------------------
def _rhs(g, data, op):
    sparse_t = torch.as_tensor(np.asarray(g.interpolate(data, ret_image=False)))
    b = sparse_t * torch.as_tensor(g.plan.pre_weights)
    return b.reshape(*b.shape[:-1], *op.natural_shape)


def test_solve_shapes_default_and_errors(grog_setup):
    from pygrog_amd.operator import MaskedFFT, SparseFFT

    g, data, sm = grog_setup["grog"], grog_setup["data"], grog_setup["smaps"]
    op = SparseFFT(plan=g.plan, smaps=sm)
    b = _rhs(g, data, op)
    x_cg = op.solve(b, method="cg", max_iter=10, damp=1e-3)
    x_ls = op.solve(b, method="lsmr", max_iter=10)
    assert x_cg.shape == tuple(grog_setup["image_shape"]) == x_ls.shape
    assert torch.isfinite(x_cg).all() and torch.isfinite(x_ls).all()
    assert op.toeplitz is True
    assert torch.equal(op.solve(b, max_iter=5), op.solve(b, method="cg", max_iter=5))
    with pytest.raises(ValueError, match="Unknown method"):
        op.solve(b, method="bogus")
    kgrid, mplan = g.interpolate(data, grid=True)
    mop = MaskedFFT(plan=mplan, smaps=sm)
    xm = mop.solve(np.asarray(kgrid), method="cg", max_iter=5, damp=1e-3)
    assert isinstance(xm, np.ndarray) and xm.shape == tuple(grog_setup["image_shape"]) and np.isfinite(xm).all()
    xl = mop.solve(torch.as_tensor(kgrid), method="lsmr", max_iter=10)
    assert xl.shape == tuple(grog_setup["image_shape"]) and torch.isfinite(xl).all()


def test_polynomial_preconditioner_on_device_operator(grog_setup):
    from pygrog_amd import PolynomialPreconditioner
    from pygrog_amd.operator import SparseFFT

    g, data, sm = grog_setup["grog"], grog_setup["data"], grog_setup["smaps"]
    op = SparseFFT(plan=g.plan, smaps=sm)
    pc = PolynomialPreconditioner(op, degree=3, n_power_iter=5)
    assert pc.degree == 3 and len(pc.coeffs) == 4 and pc.spectrum[0] == 0.0 and pc.spectrum[1] > 0.0
    x = torch.randn(*grog_setup["image_shape"], dtype=torch.complex64)
    assert pc.apply(x).shape == x.shape and torch.isfinite(pc.apply(x)).all()
    b = _rhs(g, data, op)
    xp = op.solve(b, method="cg", max_iter=10, damp=1e-3, preconditioner=PolynomialPreconditioner(op, degree=2,
                                                                                                n_power_iter=5))
    assert xp.shape == tuple(grog_setup["image_shape"]) and torch.isfinite(xp).all()
    with pytest.raises(RuntimeError, match="sample_shape"):
        PolynomialPreconditioner(SparseFFT(plan=g.plan), degree=2, n_power_iter=3)
    pc2 = PolynomialPreconditioner(SparseFFT(plan=g.plan), degree=2, n_power_iter=3,
                                   sample_shape=(3, *grog_setup["image_shape"]))
    assert pc2.spectrum[1] > 0.0


@pytest.mark.parametrize("n", [0, 1, 255, 257, 300_001, 4_000_003])
def test_fp64_vdot_kernel(device, n):
    """The CG scalars' reduction: <a, b> of complex64 vectors summed in
    float64 against numpy's complex128 dot of the same values (exact
    products, so only the summation order differs), and deterministic."""
    from pygrog_amd._solve._solvers import _vdot_c64_f64

    rng = np.random.default_rng(n)
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    b = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ref = np.vdot(a.astype(np.complex128), b.astype(np.complex128))
    ad, bd = torch.from_numpy(a).to(device), torch.from_numpy(b).to(device)
    got = _vdot_c64_f64(ad, bd).cpu().numpy()
    scale = max(1.0, float(np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64)))
    assert abs(got[0] - ref.real) <= 1e-13 * scale and abs(got[1] - ref.imag) <= 1e-13 * scale
    assert np.array_equal(_vdot_c64_f64(ad, bd).cpu().numpy(), got)
    if n > 1:
        # strided and lazily conjugated views are materialised first
        conj = _vdot_c64_f64(ad[::2].conj(), bd[::2]).cpu().numpy()
        r2 = np.vdot(np.conj(a[::2]).astype(np.complex128), b[::2].astype(np.complex128))
        assert abs(conj[0] - r2.real) <= 1e-13 * scale and abs(conj[1] - r2.imag) <= 1e-13 * scale


def test_cg_fp64_scalars_vs_oracle_cg(device):
    """SENSE CG through the Toeplitz normal operator, FP32 matvecs with
    FP64 scalars, against the oracle's float64 CG on the same system."""
    from pygrog_amd.operator import SparseFFT

    rng = np.random.default_rng(21)
    grid, image, n, C = (48, 48), (40, 40), 3000, 4
    idx = rng.integers(0, 48 * 48, size=n)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    sm = (rng.standard_normal((C, *image)) + 1j * rng.standard_normal((C, *image))).astype(np.complex64) * 0.5
    b = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    op = SparseFFT(grid, image, idx, w, smaps=torch.as_tensor(sm).to(device), device=device, toeplitz=True)
    x = op.solve(torch.as_tensor(b).to(device), method="cg", max_iter=15, damp=1e-2, tol=0.0)
    p, inv = _perm(idx)
    sqw = np.sqrt(w)[p]

    def normal(v):
        return oop.sparse_adjoint(oop.sparse_forward(v, idx[p], sqw, inv, grid, image, sm), idx[p], sqw, p, grid,
                                  image, sm)

    rhs = oop.sparse_adjoint(b, idx[p], sqw, p, grid, image, sm)
    ref = oop.cg(normal, rhs, damp=1e-2, max_iter=15, tol=0.0)
    assert _rel(_np(x), ref) <= 1e-5


def test_cg_stops_like_the_synchronous_rule(device):
    """The device CG reads each iteration's residual test one iteration
    late (no host wait per iteration) but stops where testing every
    iteration stops: the callbacks arrive in order, the first one under
    tol is the last, and the result equals a run of exactly that many
    iterations, bit for bit."""
    from pygrog_amd._solve import cg
    from pygrog_amd.operator import SparseFFT

    rng = np.random.default_rng(22)
    grid, image, n, C = (40, 40), (32, 32), 2500, 3
    idx = rng.integers(0, 40 * 40, size=n)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    sm = (rng.standard_normal((C, *image)) + 1j * rng.standard_normal((C, *image))).astype(np.complex64) * 0.5
    b = torch.as_tensor((rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64))
    op = SparseFFT(grid, image, idx, w, smaps=torch.as_tensor(sm).to(device), device=device, toeplitz=True)
    b = b.to(device)
    seen = []
    x = cg(op, b, damp=1e-2, max_iter=60, tol=3e-3, callback=lambda k, xk, rel: seen.append((k, rel, xk)))
    ks = [k for k, _, _ in seen]
    assert ks == list(range(len(seen))) and 2 < len(seen) < 60
    assert seen[-1][1] < 3e-3 and all(rel >= 3e-3 for _, rel, _ in seen[:-1])
    assert torch.equal(x, seen[-1][2])
    x_fixed = cg(op, b, damp=1e-2, max_iter=len(seen), tol=0.0)
    assert torch.equal(x, x_fixed)
    # max_iter reached without convergence: the last iterate
    x3 = cg(op, b, damp=1e-2, max_iter=2, tol=0.0)
    assert torch.equal(x3, cg(op, b, damp=1e-2, max_iter=2, tol=1e-30))


@pytest.mark.parametrize("damp", [0.0, 1e-2])
def test_cg_fused_step_equals_generic_loop(device, damp):
    """The fused device CG step (pg_cg_step_xr / pg_cg_step_p, taken without
    a preconditioner) gives the generic loop's iterates bit for bit; an
    identity preconditioner runs the generic loop with the same arithmetic
    (z = r, <r, z> = <r, r>)."""
    from pygrog_amd._solve import cg
    from pygrog_amd.operator import SparseFFT

    class _Identity:
        def apply(self, r):
            return r

    rng = np.random.default_rng(23)
    grid, image, n, C = (48, 48), (40, 40), 4000, 4
    idx = rng.integers(0, 48 * 48, size=n)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    sm = (rng.standard_normal((C, *image)) + 1j * rng.standard_normal((C, *image))).astype(np.complex64) * 0.5
    b = torch.as_tensor((rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64))
    op = SparseFFT(grid, image, idx, w, smaps=torch.as_tensor(sm).to(device), device=device, toeplitz=True)
    b = b.to(device)
    fused, generic = [], []
    x_f = cg(op, b, damp=damp, max_iter=15, tol=0.0, callback=lambda k, xk, rel: fused.append((k, rel, xk.clone())))
    x_g = cg(op, b, damp=damp, max_iter=15, tol=0.0, preconditioner=_Identity(),
             callback=lambda k, xk, rel: generic.append((k, rel, xk.clone())))
    assert torch.equal(x_f, x_g)
    assert len(fused) == len(generic) == 15
    for (kf, rf, xf), (kg, rg, xg) in zip(fused, generic):
        assert kf == kg and rf == rg and torch.equal(xf, xg)


# ---------------------------------------------------------------------------
# GrogInterpolator (reference tests/test_grog_interpolator.py)
# ---------------------------------------------------------------------------
def _make_grog(rng, shape, nc, nv, nr, image_shape=None, **kw):
    from pygrog_amd.calib import GrogInterpolator

    coords = rng.standard_normal((nv, nr, 2)).astype(np.float32)
    g = GrogInterpolator(shape=shape, coords=coords, kernel_width=2, oversamp=2.0,
                         image_shape=image_shape or shape, **kw)
    calib = (rng.standard_normal((nc, *shape)) + 1j * rng.standard_normal((nc, *shape))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=0.01, precision=1)
    return g, coords


def test_ret_image_matches_sparsefft_adjoint_rss(device):
    from pygrog_amd.operator import SparseFFT

    rng = np.random.default_rng(123)
    g, _ = _make_grog(rng, (12, 12), 3, 5, 7, image_shape=(10, 10))
    data = (rng.standard_normal((3, 5, 7)) + 1j * rng.standard_normal((3, 5, 7))).astype(np.complex64)
    sparse = g.interpolate(data, ret_image=False)
    image = g.interpolate(data, ret_image=True)
    op = SparseFFT(plan=g.plan)
    st = torch.as_tensor(np.asarray(sparse))
    exp = op.adjoint(st * g.plan.pre_weights.to(st.dtype).unsqueeze(0))
    exp = exp.abs().square().sum(0).sqrt().cpu().numpy()
    np.testing.assert_allclose(image, exp, rtol=1e-5, atol=1e-5)
    assert sparse.shape == (3, int(g.plan.n_samples)) and image.shape == (10, 10)


def test_interpolate_sparse_expansion_and_errors(device):
    from pygrog_amd.calib import GrogInterpolator

    rng = np.random.default_rng(321)
    g, _ = _make_grog(rng, (14, 14), 2, 6, 8, kernel_shape="square")
    data = (rng.standard_normal((2, 6, 8)) + 1j * rng.standard_normal((2, 6, 8))).astype(np.complex64)
    sparse = g.interpolate(data, ret_image=False)
    assert sparse.shape == (2, int(g.plan.n_samples))
    assert sparse.shape[1] != int(np.prod(g.plan.grid_shape)) and sparse.shape[1] > 48
    g2 = GrogInterpolator(shape=(10, 10), coords=rng.standard_normal((3, 5, 2)).astype(np.float32))
    with pytest.raises(RuntimeError, match="calc_interp_table"):
        g2.interpolate(data[:, :3, :5])


def test_shot_by_shot_sums_to_full_grid(device):
    rng = np.random.default_rng(99)
    g, coords = _make_grog(rng, (12, 12), 3, 4, 6)
    data = (rng.standard_normal((3, 4, 6)) + 1j * rng.standard_normal((3, 4, 6))).astype(np.complex64)
    with pytest.raises(RuntimeError, match="No shots accumulated"):
        g.collect_shots()
    for v in range(4):
        assert g(data[:, v, :], shot_index=v) is None
    shots = g.collect_shots()
    assert shots.shape == (3, *g.plan.grid_shape)
    full, _ = g.interpolate(data, grid=True)
    assert _rel(_np(shots).reshape(3, -1), np.asarray(full).reshape(3, -1)) <= 1e-5
    from oracle import plan as oplan

    p = oplan.attach_fft_plan(oplan.create_plan_single((12, 12), coords, 2.0, 2, "circle"))
    e = g._interp_idx.reshape(-1).cpu().numpy()
    ref, _ = ogrid.grid_reference(data.reshape(3, -1), p, e.reshape(p["distances"].shape[:-1]),
                                  g._interp_kernel.cpu().numpy())
    assert _rel(_np(shots).reshape(3, -1), ref) <= 1e-5
