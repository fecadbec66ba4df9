"""Radial self-calibration (train_grappa with coords) in FP64 on the device
against the oracle, and the autograd wrappers grog_measure /
grog_backproject (gradients are the adjoint operators)."""

import numpy as np
import pytest
import scipy.linalg as sla
import torch

from oracle import table as otable

pytestmark = pytest.mark.gpu


def test_radial_self_calibration_vs_oracle(device):
    from pygrog_amd.calib import train_grappa

    rng = np.random.default_rng(3)
    nc, nr, ns = 6, 16, 40
    Q = np.linalg.qr(rng.standard_normal((nc, nc)) + 1j * rng.standard_normal((nc, nc)))[0]
    A = Q @ np.diag(0.3 * rng.standard_normal(nc) + 0.2j * rng.standard_normal(nc)) @ np.linalg.inv(Q)
    B = Q @ np.diag(0.3 * rng.standard_normal(nc) + 0.2j * rng.standard_normal(nc)) @ np.linalg.inv(Q)
    th = np.pi * np.arange(nr) / nr
    coords = np.zeros((nr, ns, 2))
    calib = np.zeros((nc, nr, ns), np.complex128)
    for r in range(nr):
        d = np.array([np.cos(th[r]), np.sin(th[r])]) * 0.5
        coords[r] = (np.arange(ns) - ns // 2)[:, None] * d[None]
        G = sla.expm(d[0] * A + d[1] * B)
        y = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
        for i in range(ns):
            calib[:, r, i] = y
            y = y @ G
    noisy = calib + 1e-3 * (rng.standard_normal(calib.shape) + 1j * rng.standard_normal(calib.shape))
    for lam in (0.0, 1e-4):
        ops = train_grappa(noisy, lam, coords=coords, device=device)
        ref = otable.radial_unit_operators(noisy, coords, lam)
        for ax in ("x", "y"):
            got = ops.__dict__[ax].cpu().numpy()
            assert np.abs(got - ref[ax]).max() <= 1e-8 * np.abs(ref[ax]).max()
    exact = train_grappa(calib, None, coords=coords, device=device)
    np.testing.assert_allclose(exact.x.cpu().numpy(), sla.expm(A.T), atol=1e-8)
    np.testing.assert_allclose(exact.y.cpu().numpy(), sla.expm(B.T), atol=1e-8)
    with pytest.raises(ValueError, match="radial"):
        train_grappa(calib[:, :, :10], None, coords=coords, device=device)


def test_autograd_wrappers_give_adjoint_gradients(device):
    from pygrog_amd.interop import grog_backproject, grog_measure
    from pygrog_amd.operator import SparseFFT

    rng = np.random.default_rng(5)
    idx = rng.integers(0, 32 * 32, 700)
    w = rng.uniform(0.5, 1.5, 700).astype(np.float32)
    smaps = torch.randn(3, 24, 24, dtype=torch.complex64, device=device)
    op = SparseFFT((32, 32), (24, 24), idx, w, smaps=smaps, device=device)
    x = torch.randn(24, 24, dtype=torch.complex64, device=device, requires_grad=True)
    y = torch.randn(3, 700, dtype=torch.complex64, device=device)
    # d/dx Re<A x, y> = A^H y (Wirtinger convention of torch: grad = conj-linear dual)
    loss = torch.vdot(grog_measure(x, op).reshape(-1), y.reshape(-1)).real
    loss.backward()
    ref = op.adjoint(y)
    torch.testing.assert_close(x.grad, ref, rtol=1e-4, atol=1e-4)
    yy = y.clone().requires_grad_(True)
    z = torch.randn(24, 24, dtype=torch.complex64, device=device)
    loss2 = torch.vdot(grog_backproject(yy, op).reshape(-1), z.reshape(-1)).real
    loss2.backward()
    torch.testing.assert_close(yy.grad, op.forward(z), rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("dims,C", [((48, 48), 16), ((16, 16, 16), 16), ((14, 16, 18), 12)])
def test_cartesian_unit_operators_vs_oracle(device, dims, C):
    """Cartesian unit operators fitted on the device -- the FP64 normal-
    equation products from pg_grappa_gram_f64 and a host C x C solve for
    lamda > 0, QR for lamda = 0 -- against the oracle's LAPACK ridge fit of
    the unit-norm calibration (gelsd), on regions large enough for the
    device path, 2D and 3D (cubic and not)."""
    from pygrog_amd.calib import train_grappa

    rng = np.random.default_rng(sum(dims) + C)
    calib = rng.standard_normal((C, *dims)) + 1j * rng.standard_normal((C, *dims))
    # a smooth part so the fit is not pure noise
    grid = np.meshgrid(*[np.linspace(-1, 1, n) for n in dims], indexing="ij")
    calib += 3 * np.exp(1j * sum((k + 1) * g for k, g in enumerate(grid)))[None] * rng.standard_normal((C,) + (1,) * len(dims))
    for lam in (1e-3, 1e-2, 0.0):
        ops = train_grappa(calib, lam, device=device)
        ref = otable.unit_operators(calib, lam)
        for ax in ("x", "y", "z"):
            if ref[ax] is None:
                assert ops.__dict__[ax] is None
                continue
            got = ops.__dict__[ax].cpu().numpy()
            assert np.abs(got - ref[ax]).max() <= 1e-9 * np.abs(ref[ax]).max(), (lam, ax)
