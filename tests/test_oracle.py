"""Pinning the oracle (CPU).

* C restatement of the reference CPU ops vs golden vectors written by the
  reference's own C++ (tests/golden/make_golden.py) -- bitwise;
* the reference's known-answer tests (tests/test_fast_binning.py:101-131,
  tests/test_sparse_fft.py:45-106) on the oracle ops;
* plan: neighbourhood sizes, hand-computed cases, invariants (|d| <= 0.5,
  target = nearest + offset, sum of weights = 1 per hit cell);
* quantisation: shifts on the quantisation lattice map to centre + k, ties
  round half to even;
* FP64 table: identity at zero shift, semigroup, ridge = normal equations,
  exact recovery of a known operator, product order;
* FP64 gridding oracle vs brute force.
"""

import pathlib

import numpy as np
import pytest
import torch

import oracle
from oracle import grid as ogrid
from oracle import ops
from oracle import plan as oplan
from oracle import table as otable

GOLD = np.load(pathlib.Path(__file__).parent / "golden" / "ref_ops.npz")


@pytest.mark.parametrize("tag", ["c64", "c128"])
def test_ops_match_reference_golden(tag):
    g = GOLD
    out = ops.interpolate(g[f"interp_{tag}_data"], g[f"interp_{tag}_idx"], g[f"interp_{tag}_kernel"])
    np.testing.assert_allclose(out, g[f"interp_{tag}_out"], rtol=1e-6 if tag == "c64" else 1e-14, atol=0)
    for n in (300, 6000):
        key = f"{tag}_{n}"
        grid = g[f"sc_{key}_grid0"].copy()
        ops.scatter_add(grid, g[f"sc_{key}_data"], g[f"sc_{key}_idx"], g[f"sc_{key}_w"])
        assert np.array_equal(grid, g[f"sc_{key}_out"]), "scatter_add not bitwise equal to the reference"
        ga = ops.gather(g[f"sc_{key}_grid0"], g[f"sc_{key}_idx"], g[f"sc_{key}_w"])
        assert np.array_equal(ga, g[f"ga_{key}_out"])
        p = np.zeros(g[f"psf_{key}_out"].shape, dtype=g[f"psf_{key}_out"].dtype)
        ops.psf_scatter_scalar(p, g[f"sc_{key}_idx"], g[f"sc_{key}_w"].astype(p.dtype))
        assert np.array_equal(p, g[f"psf_{key}_out"])
    po = np.zeros_like(g[f"po_{tag}_out"])
    ops.psf_scatter_outer(po, g[f"po_{tag}_idx"], g[f"po_{tag}_w"], g[f"po_{tag}_b"])
    assert np.array_equal(po, g[f"po_{tag}_out"])
    pb = np.zeros_like(g[f"pb_{tag}_out"])
    ops.psf_scatter_outer_basis(pb, g[f"po_{tag}_idx"], g[f"po_{tag}_w"], g[f"pb_{tag}_basis"], g[f"pb_{tag}_t"])
    assert np.array_equal(pb, g[f"pb_{tag}_out"])


def test_ops_match_built_reference_when_available():
    ref = oracle.ref_ext()
    if ref is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(9)
    idx = np.sort(rng.integers(0, 77, 9000))
    w = rng.random(9000).astype(np.float32)
    d = (rng.standard_normal(9000) + 1j * rng.standard_normal(9000)).astype(np.complex64)
    a = np.zeros(77, np.complex64)
    ops.scatter_add(a, d, idx, w)
    b = torch.zeros(77, dtype=torch.complex64)
    ref.scatter_add(b, torch.from_numpy(d), torch.from_numpy(idx), torch.from_numpy(w))
    assert np.array_equal(a, b.numpy())


def test_ops_known_answers():
    g = np.zeros(10, np.complex64)
    ops.scatter_add(g, np.array([1 + 2j], np.complex64), np.array([3]), np.array([0.5], np.float32))
    assert g[3] == 0.5 + 1j and np.abs(g[:3]).max() == 0 and np.abs(g[4:]).max() == 0
    g = np.zeros(16, np.complex64)
    ops.scatter_add(g, np.ones(256, np.complex64), np.full(256, 7), np.ones(256, np.float32))
    assert g[7] == 256
    g = np.zeros(8, np.complex64)
    ops.scatter_add(g, np.ones(3, np.complex64), np.array([0, 0, 4]), np.ones(3, np.float32))
    assert g[0] == 2 and g[4] == 1 and g[1] == 0
    out = ops.gather(np.arange(8).astype(np.complex64), np.array([0, 3, 7]), np.ones(3, np.float32))
    assert np.array_equal(out, np.array([0, 3, 7], np.complex64))
    out = ops.gather(np.ones(4, np.complex64), np.array([0, 1, 2]), np.array([2, 0.5, 3], np.float32))
    assert np.array_equal(out, np.array([2, 0.5, 3], np.complex64))
    # out-of-range indices are skipped / read as zero
    g = np.zeros(4, np.complex64)
    ops.scatter_add(g, np.ones(2, np.complex64), np.array([-1, 4]), np.ones(2, np.float32))
    assert np.abs(g).max() == 0
    assert ops.gather(np.ones(4, np.complex64), np.array([5]), np.ones(1, np.float32))[0] == 0


@pytest.mark.parametrize("ndim,w,shape,kw", [
    (2, 2, "circle", 3), (2, 2, "square", 4), (2, 3, "circle", 9), (2, 3, "square", 9), (2, 3, "cross", 5),
    (3, 2, "sphere", 4), (3, 2, "cube", 8), (3, 3, "sphere", 19), (3, 3, "cube", 27), (3, 3, "cross", 7),
    (2, 1, "circle", 1), (2, 4, "square", 16),
])
def test_kernel_offset_counts(ndim, w, shape, kw):
    offs = oplan.kernel_offsets(ndim, w, shape)
    assert offs.shape == (kw, ndim)
    lo = -((w - 1) // 2)
    assert offs.min() >= lo and offs.max() <= lo + w - 1
    assert len({tuple(o) for o in offs}) == kw


def test_plan_hand_computed_2d():
    # shape 8x8, oversamp 1: amp = 2*(8-1-4) = 6 -> coords scaled so max |c| = 3;
    # origins -4, steps 1.  A sample at (1.5, -2.0) -> (3.0, -2.0) after
    # scaling by max(|x|) = 1.5 -> 3 and max(|y|) = 2 -> 3 (second sample).
    coords = np.array([[1.5, -2.0], [0.0, 2.0]], np.float32)
    p = oplan.create_plan_single((8, 8), coords, 1.0, 1, "square")
    # scaled: x: [3, 0]; y: [-3, 3] -> frac = c + 4 -> [7, 1] and [4, 7]
    assert p["grid_shape"] == (8, 8)
    assert list(p["target_idx"].reshape(-1)) == [7 * 8 + 1, 4 * 8 + 7]
    assert np.all(p["distances"] == 0)
    assert np.all(p["weights"].reshape(-1) == 1.0)


def test_plan_hand_computed_shift_and_validity():
    # one sample at x = 0.3 (after scaling), width 2 square, oversamp 1:
    # nearest 0 -> offsets {0, 1}: d = 0 - 0.3 = -0.3 (valid), 1 - 0.3 = 0.7
    # (invalid, |d| > 0.5).  y = 0 exactly.  Scale x by a second sample
    # at x = -3 (max) so that 0.3 survives unchanged (factor 1).
    coords = np.array([[0.3, 0.0], [-3.0, 3.0]], np.float32)
    p = oplan.create_plan_single((8, 8), coords, 1.0, 2, "square")
    t = p["target_idx"].reshape(2, 4)
    d = p["distances"].reshape(2, 4, 2)
    # sample 0: nearest (4, 4); offsets (0,0),(0,1),(1,0),(1,1)
    assert t[0, 0] == 4 * 8 + 4
    assert np.isclose(d[0, 0, 0], np.float32(0.0) - np.float32(0.3)) and d[0, 0, 1] == 0
    assert t[0, 1] == -1 and t[0, 2] == -1 and t[0, 3] == -1  # shifts 1 or 0.7 -> invalid
    assert t[1, 0] == 1 * 8 + 7  # x=-3 -> 1, y=3 -> 7


def _invariant_checks(p, coords_shape):
    tflat = p["target_idx"].reshape(-1)
    dist = p["distances"].reshape(-1, p["ndim"])
    valid = tflat >= 0
    assert np.all(np.abs(dist[valid]) <= 0.5)
    # target = nearest + offset: decode the flat index and compare with the shift
    gs = np.array(p["grid_shape"])
    nd = np.stack(np.unravel_index(tflat[valid], p["grid_shape"]), -1)
    assert np.all((nd >= 0) & (nd < gs))
    # sum of weights over the replicas hitting a cell is 1
    G = int(np.prod(gs))
    wsum = np.zeros(G)
    np.add.at(wsum, tflat[valid], p["weights"].reshape(-1)[valid].astype(np.float64))
    hit = wsum > 0
    np.testing.assert_allclose(wsum[hit], 1.0, rtol=2e-6)


@pytest.mark.parametrize("shape,osf,w,ks", [((32, 32), 2.0, 2, "circle"), ((20, 24), 1.5, 3, "square"),
                                            ((12, 12, 12), None, 2, "sphere"), ((10, 10, 10), 1.0, 3, "cross")])
def test_plan_invariants(shape, osf, w, ks):
    rng = np.random.default_rng(4)
    coords = rng.uniform(-0.5, 0.5, (50, 7, len(shape))).astype(np.float32)
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, osf, w, ks))
    _invariant_checks(p, coords.shape)
    assert np.all(np.diff(p["indices"]) >= 0)
    assert np.array_equal(p["sort_perm"][p["inv_perm"]], np.arange(p["n_samples"]))


def test_integer_coords_give_zero_shift_and_centre_entry():
    # samples exactly on grid points: nearest replica has d = 0 and table
    # entry (q/2)(1 + n + n^2) (3D), (q/2)(1 + n) (2D)
    rng = np.random.default_rng(1)
    pts = rng.integers(-5, 6, (40, 3)).astype(np.float32)
    pts[0] = [-6, -6, -6]  # sets the per-axis max |c| = 6 -> scale 6/6
    p = oplan.create_plan_single((13, 13, 13), pts, 1.0, 1, "cube")
    assert np.all(p["distances"] == 0)
    e = oplan.quantize(p["distances"], 41**3, quant=40)
    assert np.all(e == 20 * (1 + 41 + 41 * 41))
    p2 = oplan.create_plan_single((13, 13), pts[:, :2], 1.0, 1, "square")
    e2 = oplan.quantize(p2["distances"], 11**2, precision=1)
    assert np.all(e2 == 5 * (1 + 11))


def test_quantize_lattice_and_ties():
    # precision 1: d = k/10 -> index 5 + k; ties at 0.05 + 0.1k go to even
    d = np.array([[0.0], [0.1], [-0.3], [0.5], [-0.5], [0.04], [0.05], [0.15], [0.25], [-0.05], [-0.15]], np.float32)
    e = oplan.quantize(d, 11, precision=1)
    assert list(e) == [5, 6, 2, 10, 0, 5, 5, 7, 7, 5, 3]
    # 1/32 quantisation: d = k/32 exactly -> 16 + k
    k = np.arange(-16, 17)
    e = oplan.quantize((k / 32).astype(np.float32)[:, None], 33, quant=32)
    assert np.array_equal(e, 16 + k)
    # 2D flat index = i0 + 11 i1, 3D adds 121 i2 (axis 0 fastest in the entry)
    e = oplan.quantize(np.array([[0.2, -0.1]], np.float32), 121, precision=1)
    assert e[0] == 7 + 11 * 4
    e = oplan.quantize(np.array([[0.2, -0.1, 0.3]], np.float32), 1331, precision=1)
    assert e[0] == 7 + 11 * 4 + 121 * 8
    # clamping for out-of-window shifts
    e = oplan.quantize(np.array([[0.9, 0.9]], np.float32), 121, precision=1)
    assert e[0] == 120


def test_quant_params_reference_quirks():
    assert oplan.quant_params(0.5, 1) == (10.0, 0.1, 11)
    assert oplan.quant_params(0.5, 2)[2] == 101
    assert oplan.quant_params(0.5, 3)[2] == 1001
    assert oplan.quant_params(quant=40) == (40.0, 0.025, 41)


def _rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_ridge_equals_normal_equations():
    rng = np.random.default_rng(3)
    A, B = _rand_c(rng, 60, 5), _rand_c(rng, 60, 5)
    lam = 0.3
    X = otable.ridge(A, B, lam)
    Xn = np.linalg.solve(A.conj().T @ A + lam * np.eye(5), A.conj().T @ B)
    np.testing.assert_allclose(X, Xn, rtol=1e-10, atol=1e-12)


def test_unit_operator_exact_recovery():
    # calibration generated by a known shift operator along axis 0:
    # c[:, i+1, j] = G c[:, i, j]  ->  lamda = 0 recovers G exactly
    rng = np.random.default_rng(5)
    C = 4
    Q, _ = np.linalg.qr(_rand_c(rng, C, C))
    G = Q @ np.diag(np.exp(1j * rng.uniform(-0.3, 0.3, C)))
    calib = np.zeros((C, 9, 7), np.complex128)
    calib[:, 0, :] = _rand_c(rng, C, 7)
    for i in range(8):
        calib[:, i + 1, :] = G @ calib[:, i, :]
    ops_ = otable.unit_operators(calib, 0.0)
    np.testing.assert_allclose(ops_["x"], G, atol=1e-9)


def test_unit_operator_exact_recovery_3d_axis_labels():
    """3D calibration generated by three known commuting shift operators,
    c[:, i, j, k] = A0^i A1^j A2^k v: lamda = 0 recovers the operator of
    calibration axis 2 as "x", of axis 1 as "y" and of axis 0 as "z" (the
    reference's labelling, src/pygrog/calib/_grappa.py:347-368); a swapped
    or transposed result fails (the three operators differ and are not
    symmetric)."""
    rng = np.random.default_rng(15)
    C = 4
    Q, _ = np.linalg.qr(_rand_c(rng, C, C))
    P = Q @ np.diag(rng.uniform(0.8, 1.3, C)) @ Q.conj().T  # not unitary: A^T != A^H-based forms
    Pi = np.linalg.inv(P)
    A = [P @ np.diag(np.exp(1j * rng.uniform(-0.4, 0.4, C))) @ Pi for _ in range(3)]
    v = _rand_c(rng, C)
    n0, n1, n2 = 5, 6, 7
    calib = np.zeros((C, n0, n1, n2), np.complex128)
    for i in range(n0):
        for j in range(n1):
            for k in range(n2):
                calib[:, i, j, k] = (np.linalg.matrix_power(A[0], i) @ np.linalg.matrix_power(A[1], j)
                                     @ np.linalg.matrix_power(A[2], k) @ v)
    ops_ = otable.unit_operators(calib, 0.0)
    np.testing.assert_allclose(ops_["x"], A[2], atol=1e-8)
    np.testing.assert_allclose(ops_["y"], A[1], atol=1e-8)
    np.testing.assert_allclose(ops_["z"], A[0], atol=1e-8)
    for a in range(3):
        assert np.abs(A[a] - A[a].T).max() > 1e-3
        assert np.abs(A[a] - A[(a + 1) % 3]).max() > 1e-3


def test_table_identity_semigroup_and_order():
    rng = np.random.default_rng(6)
    calib = _rand_c(rng, 3, 10, 10)
    t, ns, nd = otable.table(calib, 0.01, 1)
    assert (ns, nd) == (11, 2) and t.shape == (121, 3, 3)
    np.testing.assert_allclose(t[5 + 11 * 5], np.eye(3), atol=1e-13)
    u = otable.unit_operators(calib, 0.01)
    # semigroup: G^{1/2} G^{1/2} = G
    h = otable.powers(u["x"], np.array([0.5], np.float32))[0]
    np.testing.assert_allclose(h @ h, u["x"], rtol=1e-9, atol=1e-11)
    # product order: table[i1*n + i0] = Gx^{d(i0)} @ Gy^{d(i1)}
    ex = otable.deltas(11, 0.1)
    gx, gy = otable.powers(u["x"], ex), otable.powers(u["y"], ex)
    np.testing.assert_allclose(t[3 + 11 * 8], gx[3] @ gy[8], rtol=1e-12, atol=1e-14)
    assert np.abs(gx[3] @ gy[8] - gy[8] @ gx[3]).max() > 1e-6  # the order matters


def test_table_3d_axis_pairing():
    rng = np.random.default_rng(7)
    calib = _rand_c(rng, 2, 6, 6, 6)
    t, ns, nd = otable.table(calib, 0.01, 1)
    u = otable.unit_operators(calib, 0.01)
    ex = otable.deltas(11, 0.1)
    gx, gy, gz = (otable.powers(u[a], ex) for a in "xyz")
    # reference pairing: entry (i0, i1, i2) = Gx[i0] Gy[i1] Gz[i2]
    np.testing.assert_allclose(t[2 + 11 * 4 + 121 * 9], gx[2] @ gy[4] @ gz[9], rtol=1e-12, atol=1e-14)
    tc, _, _ = otable.table(calib, 0.01, 1, axis_order="corrected")
    np.testing.assert_allclose(tc[2 + 11 * 4 + 121 * 9], gz[2] @ gy[4] @ gx[9], rtol=1e-12, atol=1e-14)


def test_grid_fp64_matches_brute_force():
    rng = np.random.default_rng(8)
    coords = rng.uniform(-0.5, 0.5, (6, 5, 2)).astype(np.float32)
    p = oplan.attach_fft_plan(oplan.create_plan_single((10, 10), coords, 2.0, 2, "circle"))
    C = 3
    t, _, _ = otable.table(_rand_c(rng, C, 8, 8), 0.01, 1)
    e = oplan.quantize(p["distances"], t.shape[0])
    y = _rand_c(rng, C, 30)
    out = ogrid.grid_fp64(y, p, e, t)
    brute = np.zeros((C, 400), np.complex128)
    tf = p["target_idx"].reshape(-1)
    w = p["weights"].reshape(-1)
    ee = e.reshape(-1)
    for r in range(tf.size):
        if tf[r] >= 0:
            brute[:, tf[r]] += w[r] * (t[ee[r]] @ y[:, r // 3])
    np.testing.assert_allclose(out, brute, rtol=1e-12, atol=1e-14)
    sub = ogrid.grid_fp64(y, p, e, t, cells=np.array([tf[tf >= 0][0], 7]))
    np.testing.assert_allclose(sub, brute[:, [tf[tf >= 0][0], 7]], rtol=1e-12, atol=1e-14)
    # replicas taken in several chunks, complex64 input widened per replica
    np.testing.assert_allclose(ogrid.grid_fp64(y, p, e, t, chunk=7), brute, rtol=1e-12, atol=1e-14)
    y32 = y.astype(np.complex64)
    np.testing.assert_allclose(ogrid.grid_fp64(y32, p, e, t, chunk=5),
                               ogrid.grid_fp64(y32.astype(np.complex128), p, e, t), rtol=1e-13, atol=1e-15)
    gr, dens = ogrid.grid_reference(y.astype(np.complex64), p, e, t.astype(np.complex64))
    assert ogrid.rel_l2(gr, brute) < 1e-6
    np.testing.assert_allclose(dens[dens > 0], 1.0, rtol=1e-6)
    # the cell-subset reference flow gives the full flow's columns bit for bit
    cells = np.array([tf[tf >= 0][3], 7, tf[tf >= 0][0], 399, tf[tf >= 0][-1]])
    sub = ogrid.grid_reference(y.astype(np.complex64), p, e, t.astype(np.complex64), cells=cells)
    assert sub.dtype == np.complex64 and np.array_equal(sub, gr[:, cells])


def test_grid_reference_cells_bitwise_on_hot_cells():
    """Many replicas per cell (radial centre, summation order matters in
    float32): the subset flow still equals the full flow bit for bit."""
    rng = np.random.default_rng(9)
    r = (np.arange(64) - 32) / 64
    th = np.deg2rad(111.246117975) * np.arange(40)
    coords = np.stack([np.outer(np.cos(th), r), np.outer(np.sin(th), r)], -1).astype(np.float32)
    p = oplan.attach_fft_plan(oplan.create_plan_single((32, 32), coords, None, 3, "circle"))
    C = 3
    t, _, _ = otable.table(_rand_c(rng, C, 8, 8), 0.01, 1, 16)
    e = oplan.quantize(p["distances"], t.shape[0], quant=16)
    y = _rand_c(rng, C, 40 * 64).astype(np.complex64)
    full, _ = ogrid.grid_reference(y, p, e, t.astype(np.complex64))
    cells = np.concatenate([np.arange(16 * 32 + 12, 16 * 32 + 20), rng.choice(512, 50, replace=False)])
    sub = ogrid.grid_reference(y, p, e, t.astype(np.complex64), cells=cells)
    assert np.array_equal(sub, full[:, cells])


def test_radial_unit_operators_recover_known_generators():
    """Spokes generated by y_{i+1} = y_i expm(dx_s A + dy_s B) (row vectors)
    with commuting A, B: the radial fit returns expm(A^T), expm(B^T), the
    column-vector operators (a transposed or swapped result fails)."""
    import scipy.linalg as sla

    from oracle import table as otable

    rng = np.random.default_rng(2)
    nc, nr, ns = 4, 12, 24
    Q = np.linalg.qr(rng.standard_normal((nc, nc)) + 1j * rng.standard_normal((nc, nc)))[0]
    Qi = np.linalg.inv(Q)
    A = Q @ np.diag(0.3 * rng.standard_normal(nc) + 0.2j * rng.standard_normal(nc)) @ Qi
    B = Q @ np.diag(0.3 * rng.standard_normal(nc) + 0.2j * rng.standard_normal(nc)) @ Qi
    th = np.pi * np.arange(nr) / nr
    coords = np.zeros((nr, ns, 2))
    calib = np.zeros((nc, nr, ns), np.complex128)
    for r in range(nr):
        d = np.array([np.cos(th[r]), np.sin(th[r])]) * 0.7
        coords[r] = (np.arange(ns) - ns // 2)[:, None] * d[None]
        G = sla.expm(d[0] * A + d[1] * B)
        y = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
        for i in range(ns):
            calib[:, r, i] = y
            y = y @ G
    ops = otable.radial_unit_operators(calib, coords, 0.0)
    np.testing.assert_allclose(ops["x"], sla.expm(A.T), atol=1e-8)
    np.testing.assert_allclose(ops["y"], sla.expm(B.T), atol=1e-8)


def test_grid_reference_complex128_flow():
    """The complex128 reference flow equals brute force in float64 with the
    plan's sqrt weights (pre-weight times scatter weight) to rounding."""
    rng = np.random.default_rng(10)
    coords = rng.uniform(-0.5, 0.5, (5, 6, 2)).astype(np.float32)
    p = oplan.attach_fft_plan(oplan.create_plan_single((10, 10), coords, 2.0, 2, "circle"))
    C = 2
    t, _, _ = otable.table(_rand_c(rng, C, 8, 8), 0.01, 1)
    e = oplan.quantize(p["distances"], t.shape[0])
    y = _rand_c(rng, C, 30)
    g128, _ = ogrid.grid_reference(y, p, e, t, dtype=np.complex128)
    assert g128.dtype == np.complex128
    brute = np.zeros((C, 400), np.complex128)
    sw = p["sqrt_weights"].astype(np.float64)
    for q in range(p["indices"].size):
        r = p["sort_perm"][q]
        brute[:, p["indices"][q]] += sw[q] * sw[q] * (t[e.reshape(-1)[r]] @ y[:, r // 3])
    np.testing.assert_allclose(g128, brute, rtol=1e-12, atol=1e-14)
