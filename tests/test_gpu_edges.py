"""GPU plan compiler and quantisation on lattice inputs, where rounding
decides: samples on grid nodes (zero shifts, the centre table entry),
samples half-way between nodes (round-half-even nearest node, shifts of
exactly +-0.5 at the validity bound), shifts on and next to the
quantisation ties (0.05 + 0.1 k at precision 1), and a single sample.
Target indices, shifts, sorted cells and table entries are compared bit for
bit with the oracle's float32 restatement (calib/_grog.py:1081-1234,
918-998, 382-422), and so are the permutations except for the order of
cell 0's replicas when a valid one lands there (the reference's argsort
leaves that order open); the gridded k-space within rel-L2 1e-5 of the
reference flow on the same table.
"""

import numpy as np
import pytest

from oracle import grid as ogrid
from oracle import plan as oplan

pytestmark = pytest.mark.gpu

N_AXIS = 13  # 13-point axes: coordinates of largest magnitude 6 rescale exactly


def _lattice(ndim, step, kmax, n, seed, corner=True):
    """``n`` points k * step with |k| <= kmax (k * step < 6), and -6 on every
    axis -- the per-axis maximum magnitude, so the rescale to the grid is
    x -> x -- in one point (``corner``: on the grid's first node, cell 0) or
    in one point per axis."""
    rng = np.random.default_rng(seed)
    pts = (rng.integers(-kmax, kmax + 1, (n, ndim)).astype(np.float32) * np.float32(step)).astype(np.float32)
    if corner:
        pts[0] = -6.0
    else:
        for d in range(ndim):
            pts[d, d] = -6.0
    return pts


EDGE = [
    # name, ndim, coords, kernel width, kernel shape, quantisation
    ("3d_nodes_cube1", 3, lambda: _lattice(3, 1.0, 5, 40, 1), 1, "cube", dict(quant=40)),
    ("2d_nodes_square2", 2, lambda: _lattice(2, 1.0, 5, 40, 2), 2, "square", dict(precision=1)),
    ("2d_half_nodes_w3", 2, lambda: _lattice(2, 0.5, 11, 60, 3, corner=False), 3, "circle", dict(precision=1)),
    ("3d_half_nodes_sphere", 3, lambda: _lattice(3, 0.5, 11, 60, 4), 2, "sphere", dict(quant=10)),
    ("2d_twentieths", 2, lambda: _lattice(2, 0.05, 119, 300, 5, corner=False), 2, "circle", dict(precision=1)),
    ("3d_fortieths", 3, lambda: _lattice(3, 0.025, 239, 300, 6, corner=False), 2, "cube", dict(quant=40)),
    ("2d_single_sample", 2, lambda: np.array([[0.3, -0.2]], np.float32), 2, "circle", dict(precision=1)),
]


def _calib(C, shape, seed=1):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((C, *shape)) + 1j * rng.standard_normal((C, *shape))).astype(np.complex64)


@pytest.mark.parametrize("case", EDGE, ids=[c[0] for c in EDGE])
def test_lattice_plan_entries_and_grid(device, case):
    from pygrog_amd.calib import GrogInterpolator

    name, nd, mk, w, ks, q = case
    coords = mk()
    shape = (N_AXIS,) * nd
    print(f"edge case {name}: {coords.shape[0]} samples, width {w} {ks}, {q}", flush=True)
    g = GrogInterpolator(shape, coords, oversamp=1.0, kernel_width=w, kernel_shape=ks, device=device)
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, 1.0, w, ks))
    t_gpu = g.plan.target_idx.reshape(-1).cpu().numpy()
    np.testing.assert_array_equal(t_gpu, p["target_idx"].reshape(-1))
    d_gpu = g.plan.distances.cpu().numpy().reshape(p["distances"].shape)
    assert np.array_equal(d_gpu.view(np.uint32), p["distances"].view(np.uint32)), "shifts not bit-exact"
    np.testing.assert_array_equal(g.plan.indices.cpu().numpy(), p["indices"])
    np.testing.assert_allclose(g.plan.weights.reshape(-1).cpu().numpy(), p["weights"].reshape(-1), rtol=1e-6, atol=1e-7)
    valid = p["target_idx"].reshape(-1) >= 0
    assert valid.any()
    sp, ip = g.plan.sort_perm.cpu().numpy(), g.plan.inv_perm.cpu().numpy()
    sw = g.plan.sqrt_weights.cpu().numpy()
    if not np.any(p["target_idx"].reshape(-1) == 0):
        np.testing.assert_array_equal(sp, p["sort_perm"])
        np.testing.assert_array_equal(ip, p["inv_perm"])
    else:
        # a valid replica in cell 0 (the grid's first corner): the reference
        # clamps invalid replicas to cell 0 and sorts with torch.argsort
        # (stable=False), so the order among cell 0's replicas is not fixed;
        # the plan puts the invalid ones first and keeps every cell's valid
        # replicas in sample order.  Same cells, same members, a consistent
        # permutation, weights following their replicas.
        clamped = np.where(valid, p["target_idx"].reshape(-1), 0)
        n = clamped.size
        assert np.array_equal(np.sort(sp), np.arange(n))
        assert np.array_equal(clamped[sp], p["indices"])
        assert np.array_equal(ip[sp], np.arange(n))
        k0 = int(np.count_nonzero(p["indices"] == 0))
        assert np.array_equal(np.sort(sp[:k0]), np.sort(p["sort_perm"][:k0]))
        np.testing.assert_array_equal(sp[k0:], p["sort_perm"][k0:])
        inv = np.flatnonzero(~valid)
        assert np.array_equal(sp[:inv.size], inv), "invalid replicas first, in sample order"
        by_pos = np.empty(n, np.float32)
        by_pos[p["sort_perm"]] = p["sqrt_weights"]
        np.testing.assert_allclose(sw, by_pos[sp], rtol=1e-6, atol=1e-7)

    C = 3
    g.calc_interp_table(_calib(C, (8,) * nd), lamda=0.01, **q)
    table = g._interp_kernel.cpu().numpy()
    e_ref = oplan.quantize(p["distances"], table.shape[0], precision=q.get("precision", 1), quant=q.get("quant"))
    e_gpu = g._interp_idx.reshape(-1).cpu().numpy()
    np.testing.assert_array_equal(e_gpu, e_ref.reshape(-1))
    if "_nodes_" in name and "half" not in name:
        # on a node every valid replica is the node itself: zero shift, and
        # the centre entry (q/2)(1 + n + n^2) (3D) or (q/2)(1 + n) (2D)
        n = table.shape[0] ** (1.0 / nd)
        n = int(round(n))
        centre = (n // 2) * sum(n ** k for k in range(nd))
        assert np.all(p["distances"].reshape(-1, nd)[valid] == 0)
        assert np.all(e_gpu[valid] == centre)
    if "half_nodes" in name:
        # shifts of exactly 0.5 grid steps sit on the validity bound
        assert np.any(np.abs(p["distances"].reshape(-1, nd)[valid]) == 0.5)

    rng = np.random.default_rng(11)
    y = (rng.standard_normal((C, coords.shape[0])) + 1j * rng.standard_normal((C, coords.shape[0]))).astype(np.complex64)
    kgrid, _ = g.interpolate(y, grid=True)
    gref, _ = ogrid.grid_reference(y, p, e_ref, table)
    rel = ogrid.rel_l2(np.asarray(kgrid).reshape(C, -1), gref)
    assert rel <= 1e-5, f"fp32 grid rel-L2 vs reference flow {rel:.2e}"
