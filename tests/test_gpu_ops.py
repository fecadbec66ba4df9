"""The seven reference native ops on the GPU vs the C oracle and the
reference's own known-answer tests (tests/test_fast_binning.py,
tests/test_sparse_fft.py:45-106, tests/test_toeplitz_kernels.py)."""

import numpy as np
import pytest
import torch

from oracle import ops as orc
from pygrog_amd.operator import gather, scatter_add
from pygrog_amd.operator._sparse_fft import _get_torch_ext, _scatter_add

pytestmark = pytest.mark.gpu


def _sorted(n, G, device, seed=0):
    g = torch.Generator().manual_seed(seed)
    idx = torch.randint(0, G, (n,), generator=g).sort().values
    w = torch.rand(n, generator=g)
    d = torch.randn(n, dtype=torch.complex64, generator=g)
    return idx.to(device), w.to(device), d.to(device)


def test_extension_loads():
    ext = _get_torch_ext()
    for name in ("grog_interpolate", "scatter_add", "scatter_add_binned", "gather", "psf_scatter_scalar",
                 "psf_scatter_outer", "psf_scatter_outer_basis"):
        assert hasattr(ext, name)


@pytest.mark.parametrize("n,G", [(512, 64), (20000, 300), (100000, 7)])
def test_scatter_add_vs_oracle(device, n, G):
    idx, w, d = _sorted(n, G, device, seed=n)
    grid = torch.zeros(G, dtype=torch.complex64, device=device)
    scatter_add(grid, d, idx, w)
    ref = np.zeros(G, np.complex64)
    orc.scatter_add(ref, d.cpu().numpy(), idx.cpu().numpy(), w.cpu().numpy())
    np.testing.assert_allclose(grid.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)


def test_scatter_add_unsorted_and_out_of_range(device):
    g = torch.Generator().manual_seed(3)
    idx = torch.randint(-5, 70, (3000,), generator=g)
    w = torch.rand(3000, generator=g)
    d = torch.randn(3000, dtype=torch.complex64, generator=g)
    grid = torch.zeros(64, dtype=torch.complex64, device=device)
    scatter_add(grid, d.to(device), idx.to(device), w.to(device))
    ref = np.zeros(64, np.complex64)
    orc.scatter_add(ref, d.numpy(), idx.numpy(), w.numpy())
    np.testing.assert_allclose(grid.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


def test_scatter_add_accumulates_into_existing(device):
    idx, w, d = _sorted(128, 32, device, seed=1)
    grid = torch.ones(32, dtype=torch.complex64, device=device)
    ref = grid.clone()
    scatter_add(grid, d, idx, w)
    ref.index_add_(0, idx, d * w.to(torch.complex64))
    torch.testing.assert_close(grid, ref, rtol=1e-4, atol=1e-4)


def test_scatter_add_known_answers(device):
    # reference tests/test_fast_binning.py:101-131, test_sparse_fft.py:45-75
    grid = torch.zeros(10, dtype=torch.complex64, device=device)
    scatter_add(grid, torch.tensor([1 + 2j], dtype=torch.complex64, device=device),
                torch.tensor([3], device=device), torch.tensor([0.5], device=device))
    assert abs(grid[3].item() - (0.5 + 1j)) < 1e-5
    assert grid[:3].abs().max() == 0 and grid[4:].abs().max() == 0
    grid = torch.zeros(16, dtype=torch.complex64, device=device)
    scatter_add(grid, torch.ones(256, dtype=torch.complex64, device=device), torch.full((256,), 7, device=device),
                torch.ones(256, device=device))
    assert abs(grid[7].real.item() - 256) < 1e-3 and grid[7].imag.abs().item() < 1e-5
    grid = torch.zeros(8, dtype=torch.complex64, device=device)
    scatter_add(grid, torch.ones(3, dtype=torch.complex64, device=device), torch.tensor([0, 0, 4], device=device),
                torch.ones(3, device=device))
    assert grid[0].real.item() == pytest.approx(2.0) and grid[4].real.item() == pytest.approx(1.0)
    grid = torch.zeros(16, dtype=torch.complex64, device=device)
    scatter_add(grid, torch.empty(0, dtype=torch.complex64, device=device),
                torch.empty(0, dtype=torch.int64, device=device), torch.empty(0, device=device))
    assert grid.abs().max() == 0


def test_scatter_binned_matches_simple(device):
    idx, w, d = _sorted(4096, 1024, device, seed=5)
    g1 = torch.zeros(1024, dtype=torch.complex64, device=device)
    _scatter_add(g1, d, idx, w)
    edges = torch.arange(5, device=device) * 256
    bin_starts = torch.searchsorted(idx, edges)
    g2 = torch.zeros(1024, dtype=torch.complex64, device=device)
    _scatter_add(g2, d, idx, w, bin_starts=bin_starts, bin_size=256)
    torch.testing.assert_close(g1, g2, rtol=1e-4, atol=1e-4)


def test_scatter_add_c128(device):
    idx, w, d = _sorted(5000, 100, device, seed=9)
    grid = torch.zeros(100, dtype=torch.complex128, device=device)
    scatter_add(grid, d.to(torch.complex128), idx, w.double())
    ref = np.zeros(100, np.complex128)
    orc.scatter_add(ref, d.cpu().numpy().astype(np.complex128), idx.cpu().numpy(), w.cpu().numpy().astype(np.float64))
    np.testing.assert_allclose(grid.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)


def test_gather_vs_oracle_and_edges(device):
    idx, w, _ = _sorted(2000, 64, device, seed=2)
    grid = torch.randn(64, dtype=torch.complex64, device=device)
    out = gather(grid, idx, w)
    ref = orc.gather(grid.cpu().numpy(), idx.cpu().numpy(), w.cpu().numpy())
    assert np.array_equal(out.cpu().numpy(), ref)
    assert gather(grid, torch.empty(0, dtype=torch.int64, device=device), torch.empty(0, device=device)).numel() == 0
    z = gather(torch.ones(8, dtype=torch.complex64, device=device) * 99, torch.arange(8, device=device),
               torch.zeros(8, device=device))
    assert z.abs().max() < 1e-6
    # out of range reads as 0 (reference sparse_ops_cpu_impl.inl:158-163)
    o = gather(grid, torch.tensor([-1, 64, 3], device=device), torch.ones(3, device=device))
    assert o[0] == 0 and o[1] == 0 and o[2] == grid[3]


def test_scatter_gather_round_trip(device):
    data = torch.randn(32, dtype=torch.complex64, device=device)
    idx = torch.arange(32, device=device)
    w = torch.ones(32, device=device)
    grid = torch.zeros(32, dtype=torch.complex64, device=device)
    scatter_add(grid, data, idx, w)
    torch.testing.assert_close(gather(grid, idx, w), data, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("N,B,C,K", [(1000, 1, 8, 50), (333, 3, 5, 7), (64, 2, 32, 9)])
def test_grog_interpolate_vs_oracle(device, N, B, C, K):
    rng = np.random.default_rng(N + C)
    d = (rng.standard_normal((N, B, C)) + 1j * rng.standard_normal((N, B, C))).astype(np.complex64)
    k = (rng.standard_normal((K, C, C)) + 1j * rng.standard_normal((K, C, C))).astype(np.complex64)
    idx = rng.integers(0, K, N)
    out = _get_torch_ext().grog_interpolate(torch.from_numpy(d).to(device), torch.from_numpy(idx).to(device),
                                            torch.from_numpy(k).to(device))
    ref = orc.interpolate(d, idx, k)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


def test_grog_interpolate_checks(device):
    ext = _get_torch_ext()
    with pytest.raises(RuntimeError, match="3-D"):
        ext.grog_interpolate(torch.zeros(3, 2, dtype=torch.complex64, device=device),
                             torch.zeros(3, dtype=torch.int64, device=device),
                             torch.zeros(1, 2, 2, dtype=torch.complex64, device=device))
    with pytest.raises(RuntimeError, match="int64"):
        ext.grog_interpolate(torch.zeros(3, 1, 2, dtype=torch.complex64, device=device),
                             torch.zeros(3, dtype=torch.int32, device=device),
                             torch.zeros(1, 2, 2, dtype=torch.complex64, device=device))


def test_psf_scatter_scalar(device):
    idx, w, _ = _sorted(200, 64, device, seed=0)
    psf = torch.zeros(64, device=device)
    _get_torch_ext().psf_scatter_scalar(psf, idx, w + 0.1)
    ref = np.zeros(64, np.float32)
    orc.psf_scatter_scalar(ref, idx.cpu().numpy(), (w + 0.1).cpu().numpy())
    np.testing.assert_allclose(psf.cpu().numpy(), ref, rtol=1e-6, atol=1e-6)
    psf64 = torch.zeros(64, dtype=torch.float64, device=device)
    _get_torch_ext().psf_scatter_scalar(psf64, idx, (w + 0.1).double())
    ref64 = torch.zeros(64, dtype=torch.float64).index_add_(0, idx.cpu(), (w + 0.1).double().cpu())
    np.testing.assert_allclose(psf64.cpu().numpy(), ref64.numpy(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("M", [1, 3, 5])
def test_psf_scatter_outer(device, M):
    idx, w, _ = _sorted(150, 32, device, seed=M)
    b = torch.randn(150, M, dtype=torch.complex64, device=device)
    psf = torch.zeros(32, M, M, dtype=torch.complex64, device=device)
    _get_torch_ext().psf_scatter_outer(psf, idx, w + 0.1, b)
    ref = np.zeros((32, M, M), np.complex64)
    orc.psf_scatter_outer(ref, idx.cpu().numpy(), (w + 0.1).cpu().numpy(), b.cpu().numpy())
    np.testing.assert_allclose(psf.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)
    torch.testing.assert_close(psf, psf.conj().transpose(-2, -1), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("K,T", [(1, 8), (3, 16), (5, 24)])
def test_psf_scatter_outer_basis(device, K, T):
    idx, w, _ = _sorted(200, 32, device, seed=K * 100 + T)
    basis = torch.randn(K, T, dtype=torch.complex64, device=device)
    ti = torch.randint(0, T, (200,), device=device)
    psf = torch.zeros(32, K, K, dtype=torch.complex64, device=device)
    _get_torch_ext().psf_scatter_outer_basis(psf, idx, w + 0.1, basis, ti)
    ref = np.zeros((32, K, K), np.complex64)
    orc.psf_scatter_outer_basis(ref, idx.cpu().numpy(), (w + 0.1).cpu().numpy(), basis.cpu().numpy(), ti.cpu().numpy())
    np.testing.assert_allclose(psf.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


def test_cpu_tensor_rejected():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    with pytest.raises(RuntimeError, match="GPU only"):
        scatter_add(torch.zeros(4, dtype=torch.complex64), torch.ones(1, dtype=torch.complex64),
                    torch.tensor([0]), torch.ones(1))


@pytest.mark.parametrize("B,S,C,n,G", [(2, 3, 4, 500, 300), (1, 1, 1, 17, 9), (3, 5, 2, 0, 8)])
def test_gather_perm_stack_equals_per_element(device, B, S, C, n, G):
    """One launch over (batch, stack element, coil) rows gives, bit for bit,
    the per-element gathers; both equal a numpy restatement of
    out[m] = w[inv[m]] * grid[idx[inv[m]]] (zero for indices outside the grid)."""
    from pygrog_amd import _lib

    rng = np.random.default_rng(B * 100 + S * 10 + C)
    idx = rng.integers(-2, G + 2, (S, n)).astype(np.int64)
    w = rng.random((S, n), dtype=np.float32)
    inv = np.stack([rng.permutation(n) for _ in range(S)]).astype(np.int64).reshape(S, n)
    kg = (rng.standard_normal((B, S, C, G)) + 1j * rng.standard_normal((B, S, C, G))).astype(np.complex64)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    kg_d, idx_d, w_d, inv_d = t(kg), t(idx), t(w), t(inv)
    st = _lib.stream(device)
    out = torch.empty(B, S, C, n, dtype=torch.complex64, device=device)
    _lib.call("pg_gather_perm_stack_c64", _lib.ptr(kg_d), G, _lib.ptr(idx_d), _lib.ptr(w_d), _lib.ptr(inv_d), n, C, S,
              B, _lib.ptr(out), st)
    each = torch.empty_like(out)
    for b in range(B):
        for s in range(S):
            _lib.call("pg_gather_perm_c64", kg_d[b, s].data_ptr(), G, idx_d[s].data_ptr(), w_d[s].data_ptr(),
                      inv_d[s].data_ptr(), n, C, each[b, s].data_ptr(), st)
    got, ref_each = out.cpu().numpy(), each.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), ref_each.view(np.uint32))
    exp = np.zeros((B, S, C, n), np.complex64)
    for s in range(S):
        i = inv[s]
        k = idx[s][i]
        ok = (k >= 0) & (k < G)
        g = kg[:, s][:, :, np.where(ok, k, 0)]
        wi = w[s][i]
        re = (wi * g.real).astype(np.float32)
        im = (wi * g.imag).astype(np.float32)
        exp[:, s] = np.where(ok, re + 1j * im, 0).astype(np.complex64)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("B,S,C,N,kw,K", [(2, 3, 4, 50, 3, 20), (1, 1, 2, 7, 1, 5), (2, 2, 3, 0, 2, 4)])
def test_sparse_apply_stack_equals_per_element(device, B, S, C, N, kw, K):
    """One launch over every (batch, stack element) block gives, bit for bit,
    the per-element operator applies (entries outside the table give 0)."""
    from pygrog_amd import _lib

    rng = np.random.default_rng(B * 100 + S * 10 + C + N)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    y = t((rng.standard_normal((B, S, C, N)) + 1j * rng.standard_normal((B, S, C, N))).astype(np.complex64))
    idx = t(rng.integers(-1, K + 1, (S, N * kw)).astype(np.int64))
    tab = t((rng.standard_normal((K, C, C)) + 1j * rng.standard_normal((K, C, C))).astype(np.complex64))
    st = _lib.stream(device)
    out = torch.empty(B, S, C, N * kw, dtype=torch.complex64, device=device)
    _lib.call("pg_sparse_apply_stack_c64", _lib.ptr(y), _lib.ptr(idx), _lib.ptr(tab), _lib.ptr(out), B, S, C, N, kw,
              K, st)
    for s in range(S):
        ys = y[:, s].contiguous()
        each = torch.empty(B, C, N * kw, dtype=torch.complex64, device=device)
        _lib.call("pg_sparse_apply_c64", _lib.ptr(ys), idx[s].data_ptr(), _lib.ptr(tab), _lib.ptr(each), B, C, N, kw,
                  K, st)
        assert torch.equal(out[:, s], each), s
    if N:
        e = idx.cpu().numpy()
        ref = np.zeros((B, S, C, N * kw), np.complex128)
        yh, th = y.cpu().numpy().astype(np.complex128), tab.cpu().numpy().astype(np.complex128)
        for s in range(S):
            for m in range(N * kw):
                if 0 <= e[s, m] < K:
                    ref[:, s, :, m] = yh[:, s, :, m // kw] @ th[e[s, m]].T
        got = out.cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-5 * max(np.abs(ref).max(), 1.0)


@pytest.mark.parametrize("rows,n,G", [(5, 3000, 257), (1, 10, 4), (3, 0, 8)])
def test_scatter_add_rows_equals_per_row(device, rows, n, G):
    """Several grids sharing indices and weights in one launch: each row as
    pg_scatter_add_c64 on its own (atomic adds, so equal to float32
    rounding) and as a float64 np.add.at (indices outside the grid skipped)."""
    from pygrog_amd import _lib

    rng = np.random.default_rng(rows * 7 + n)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    d = (rng.standard_normal((rows, n)) + 1j * rng.standard_normal((rows, n))).astype(np.complex64)
    idx = rng.integers(-3, G + 3, n).astype(np.int64)
    w = rng.random(n, dtype=np.float32)
    d_d, idx_d, w_d = t(d), t(idx), t(w)
    st = _lib.stream(device)
    got = torch.zeros(rows, G, dtype=torch.complex64, device=device)
    _lib.call("pg_scatter_add_rows_c64", _lib.ptr(got), G, _lib.ptr(d_d), _lib.ptr(idx_d), _lib.ptr(w_d), n, rows, st)
    each = torch.zeros(rows, G, dtype=torch.complex64, device=device)
    for r in range(rows):
        _lib.call("pg_scatter_add_c64", each[r].data_ptr(), G, d_d[r].data_ptr(), _lib.ptr(idx_d), _lib.ptr(w_d), n, st)
    g, e = got.cpu().numpy(), each.cpu().numpy()
    scale = max(np.abs(e).max() if e.size else 0.0, 1.0)
    assert np.abs(g - e).max() <= 1e-5 * scale
    ref = np.zeros((rows, G), np.complex128)
    ok = (idx >= 0) & (idx < G)
    for r in range(rows):
        np.add.at(ref[r], idx[ok], (w[ok].astype(np.float64) * d[r][ok]))
    assert np.abs(g - ref).max() <= 1e-5 * scale
