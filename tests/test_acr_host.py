"""extract_acr (host-side index selection) against known answers, following
the reference's semantics (src/pygrog/calib/_acr.py:12-129)."""

import numpy as np
import pytest
import torch

from pygrog_amd.calib import extract_acr


def test_cartesian_centre_crop_and_mask():
    data = np.arange(2 * 3 * 32 * 30, dtype=np.float64).reshape(2, 3, 32, 30).astype(np.complex64)
    out = extract_acr(data, cal_width=20, ndim=2)
    assert isinstance(out, np.ndarray)
    np.testing.assert_array_equal(out, data[..., 6:26, 5:25])  # starts (32-20)//2, (30-20)//2
    mask = np.zeros((1, 32, 30), bool)
    mask[..., ::2, :] = True
    d, m = extract_acr(torch.as_tensor(data), cal_width=20, ndim=2, mask=mask)
    assert isinstance(d, torch.Tensor)
    np.testing.assert_array_equal(m.numpy(), mask[..., 6:26, 5:25])
    with pytest.raises(ValueError, match="spatial dimensions"):
        extract_acr(data)


def test_noncartesian_radial_keeps_central_readout():
    nv, nr = 7, 64
    r = (np.arange(nr) - nr // 2) / nr
    th = np.linspace(0, np.pi, nv, endpoint=False)
    coords = np.stack([np.outer(np.cos(th), r), np.outer(np.sin(th), r)], -1)[None].astype(np.float32)
    data = np.random.default_rng(0).standard_normal((1, 3, nv, nr)).astype(np.complex64)
    w = np.ones((1, 1, nv, nr), np.float32)
    d, c, ww = extract_acr(data, cal_width=16, coords=coords, weights=w, shape=128)
    # coords rescaled to +-64 by their maximum (0.5): sample j sits at radius
    # |j - 32| * 2 grid units; radius <= 8 keeps j = 28..36 on every view
    keep = np.abs(np.arange(nr) - nr // 2) * 2 <= 8
    assert d.shape == (1, 3, nv, keep.sum()) and c.shape == (1, nv, keep.sum(), 2) and ww.shape == (1, 1, nv, keep.sum())
    np.testing.assert_array_equal(d, data[..., keep])
    np.testing.assert_allclose(np.sqrt((c ** 2).sum(-1)).max(), 8.0, rtol=1e-6)
    with pytest.raises(ValueError, match="matrix size"):
        extract_acr(data, coords=coords)
