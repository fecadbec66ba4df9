"""The torch custom-op layer on the HIP device: ``torch.ops.pygrog_amd``
against the C oracle and the reference's golden vectors, torch.library's
opcheck (schema, fake kernel, functionalisation), and a torch.compile'd
function (aot_eager: traced through the fake kernels and functionalised
in-place ops) that gives the eager result bit for bit."""

import pathlib

import numpy as np
import pytest
import torch

from oracle import ops as orc

pytestmark = pytest.mark.gpu

GOLD = np.load(pathlib.Path(__file__).parent / "golden" / "ref_ops.npz")


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def test_ops_vs_oracle_and_golden(device):
    from pygrog_amd import _lib

    ns = _lib.get_torch_ext()
    assert ns.scatter_add is torch.ops.pygrog_amd.scatter_add
    g = GOLD
    key = "c64_6000"
    grid = _t(g[f"sc_{key}_grid0"], device)
    torch.ops.pygrog_amd.scatter_add(grid, _t(g[f"sc_{key}_data"], device), _t(g[f"sc_{key}_idx"], device),
                                     _t(g[f"sc_{key}_w"], device))
    assert np.array_equal(grid.cpu().numpy(), g[f"sc_{key}_out"])
    out = torch.ops.pygrog_amd.gather(_t(g[f"sc_{key}_grid0"], device), _t(g[f"sc_{key}_idx"], device),
                                      _t(g[f"sc_{key}_w"], device))
    assert np.array_equal(out.cpu().numpy(), g[f"ga_{key}_out"])
    rng = np.random.default_rng(1)
    N, B, C, K = 500, 2, 6, 11
    d = (rng.standard_normal((N, B, C)) + 1j * rng.standard_normal((N, B, C))).astype(np.complex64)
    k = (rng.standard_normal((K, C, C)) + 1j * rng.standard_normal((K, C, C))).astype(np.complex64)
    idx = rng.integers(0, K, N)
    res = torch.ops.pygrog_amd.grog_interpolate(_t(d, device), _t(idx, device), _t(k, device))
    # the library's interpolate rounds each product and sum as the C loops do
    assert np.array_equal(res.cpu().numpy(), orc.interpolate(d, idx, k))
    with pytest.raises(RuntimeError, match="dtype"):
        torch.ops.pygrog_amd.scatter_add(torch.zeros(4, dtype=torch.complex64, device=device),
                                         torch.ones(2, dtype=torch.complex128, device=device),
                                         torch.zeros(2, dtype=torch.int64, device=device),
                                         torch.ones(2, dtype=torch.float64, device=device))


def test_opcheck(device):
    rng = np.random.default_rng(2)
    n, G = 300, 40
    idx = _t(np.sort(rng.integers(0, G, n)), device)
    w = torch.rand(n, device=device)
    d = torch.randn(n, dtype=torch.complex64, device=device)
    tests = ("test_schema", "test_faketensor", "test_aot_dispatch_dynamic")
    torch.library.opcheck(torch.ops.pygrog_amd.scatter_add.default,
                          (torch.zeros(G, dtype=torch.complex64, device=device), d, idx, w), test_utils=tests)
    torch.library.opcheck(torch.ops.pygrog_amd.gather.default,
                          (torch.randn(G, dtype=torch.complex64, device=device), idx, w), test_utils=tests)
    torch.library.opcheck(torch.ops.pygrog_amd.psf_scatter_scalar.default,
                          (torch.zeros(G, device=device), idx, w), test_utils=tests)
    kern = torch.randn(5, 3, 3, dtype=torch.complex64, device=device)
    torch.library.opcheck(torch.ops.pygrog_amd.grog_interpolate.default,
                          (torch.randn(n, 2, 3, dtype=torch.complex64, device=device),
                           torch.randint(0, 5, (n,), device=device), kern), test_utils=tests)


def test_compiled_graph_equals_eager(device):
    rng = np.random.default_rng(3)
    n, G, C, K = 2000, 128, 4, 9
    idx = _t(np.sort(rng.integers(0, G, n)), device)
    w = torch.rand(n, device=device)
    k = torch.randn(K, C, C, dtype=torch.complex64, device=device)
    e = torch.randint(0, K, (n,), device=device)
    y = torch.randn(n, 1, C, dtype=torch.complex64, device=device)

    def grid_and_back(y):
        v = torch.ops.pygrog_amd.grog_interpolate(y, e, k)[:, 0, 0].contiguous()
        grid = torch.zeros(G, dtype=torch.complex64, device=y.device)
        torch.ops.pygrog_amd.scatter_add(grid, v, idx, w)
        return grid, torch.ops.pygrog_amd.gather(grid, idx, w)

    eager = grid_and_back(y)
    compiled = torch.compile(grid_and_back, backend="aot_eager", fullgraph=True)(y)
    assert torch.equal(eager[0], compiled[0]) and torch.equal(eager[1], compiled[1])
