"""complex128 data through GrogInterpolator.interpolate: computed in
complex128 with the FP64 table end to end (the reference keeps the data's
precision: csrc/torch/grog_interp.cpp:112-149, calib/_grappa.py:62-72),
never narrowed to complex64.  Checked against the oracle's complex128
reference flow (the C ops in double, the plan's sqrt weights)."""

import numpy as np
import pytest
import torch

from oracle import grid as ogrid
from oracle import plan as oplan
from oracle import table as otable

pytestmark = pytest.mark.gpu


def _setup(device, C=4, shape=(40, 40)):
    from pygrog_amd import synth
    from pygrog_amd.calib import GrogInterpolator

    coords = synth.radial_2d(12, 64, offset=0.3)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    rng = np.random.default_rng(4)
    calib = rng.standard_normal((C, 12, 12)) + 1j * rng.standard_normal((C, 12, 12))
    g.calc_interp_table(calib, lamda=1e-3, quant=16)
    y = rng.standard_normal((C, 12, 64)) + 1j * rng.standard_normal((C, 12, 64))
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, None, 3, "circle"))
    e = g._interp_idx.reshape(p["distances"].shape[:-1]).cpu().numpy()
    return g, calib, y, p, e


def test_interpolate_complex128_sparse_grid_image(device):
    g, calib, y, p, e = _setup(device)
    C = y.shape[0]
    T = g._table64().cpu().numpy()
    assert T.dtype == np.complex128
    t64, _, _ = otable.table(calib, 1e-3, 1, 16)
    assert np.abs(T - t64).max() <= 1e-9 * np.abs(t64).max()
    # sparse: the reference's C op in double, bit for bit with the same table
    sp = g.interpolate(y)
    assert sp.dtype == np.complex128
    ref_sp = ogrid.sparse_reference(y.reshape(C, -1), e, T, 9, np.complex128)
    assert np.array_equal(np.asarray(sp).reshape(C, -1), ref_sp)
    # grid: the complex128 reference flow (pre-weight, sorted order-fixed scatter)
    kg, mplan = g.interpolate(y, grid=True)
    assert kg.dtype == np.complex128
    # (the oracle plan's float32 weights come from numpy's exp: equal to the
    # device plan's to float32 rounding, so the arithmetic is compared on the
    # device plan's weights and order, the plan itself at tolerance)
    pd = dict(p)
    for k in ("indices", "sqrt_weights", "sort_perm", "inv_perm"):
        pd[k] = g.plan._dev[k].cpu().numpy()
    assert np.array_equal(pd["indices"], p["indices"])
    np.testing.assert_allclose(pd["sqrt_weights"], p["sqrt_weights"], rtol=1e-6, atol=1e-7)
    ref_g, dens = ogrid.grid_reference(y.reshape(C, -1), pd, e, T, dtype=np.complex128)
    assert ogrid.rel_l2(kg.reshape(C, -1), ref_g) <= 1e-14
    ref_o, _ = ogrid.grid_reference(y.reshape(C, -1), p, e, t64, dtype=np.complex128)
    assert ogrid.rel_l2(kg.reshape(C, -1), ref_o) <= 1e-6
    np.testing.assert_allclose(np.asarray(mplan.density).reshape(-1), dens, rtol=1e-6, atol=1e-7)
    # the complex64 path of the same data differs at float32 rounding only
    k64, _ = g.interpolate(y.astype(np.complex64), grid=True)
    assert k64.dtype == np.complex64
    assert 1e-9 < ogrid.rel_l2(k64.reshape(C, -1), ref_g) <= 1e-5
    # image: centred IFFT, crop, root-sum-of-squares in double
    img = g.interpolate(torch.from_numpy(y).to(device), ret_image=True)
    assert img.dtype == torch.float64
    ref_i = ogrid.image_rss(ref_g, tuple(g.plan.grid_shape), tuple(g.plan.image_shape))
    assert ogrid.rel_l2(img.cpu().numpy(), ref_i) <= 1e-13
    for mode in ("bf16x3", "bf16x3c"):
        with pytest.raises(ValueError, match="complex128"):
            g.interpolate(y, grid=True, mode=mode)


def test_shot_gridding_complex128(device):
    g, calib, y, p, e = _setup(device)
    C = y.shape[0]
    T = g._table64().cpu().numpy()
    for v in range(12):
        g(y[:, v, :], shot_index=v)
    shots = g.collect_shots()
    assert shots.dtype == torch.complex128
    # reference shot flow (_grog.py:697-731) in double: operator, times the
    # weight, summed per cell in natural order
    tidx = g.plan.target_idx.cpu().numpy()
    w = g.plan.weights.cpu().numpy().astype(np.float64)
    ref = np.zeros((C, int(np.prod(g.plan.grid_shape))), np.complex128)
    for v in range(12):
        t1, w1, e1 = tidx[v].reshape(-1), w[v].reshape(-1), e[v].reshape(-1)
        for r in range(t1.size):
            if t1[r] >= 0:
                ref[:, t1[r]] += (T[e1[r]] @ y[:, v, r // 9]) * w1[r]
    assert ogrid.rel_l2(shots.reshape(C, -1).cpu().numpy(), ref) <= 1e-13
