"""The centring identities the fused operators rely on (CPU): for even
lengths, fftc(x) = (-1)^(k-M) fftn((-1)^n x) and ifftc(x) = (-1)^(n-M)
ifftn((-1)^k x), with the sign helpers the operators use."""

import numpy as np
import pytest
import torch

from pygrog_amd.operator._centred import cell_signs, even_grid, grid_signs, half_sign


@pytest.mark.parametrize("shape", [(8,), (6, 10), (4, 8, 6)])
def test_centred_transforms_as_sign_modulations(shape, rng):
    x = torch.from_numpy(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    ax = tuple(range(len(shape)))
    s = grid_signs(shape, "cpu").to(torch.float64)
    h = half_sign(shape)
    fc = torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(x, dim=ax), dim=ax, norm="ortho"), dim=ax)
    torch.testing.assert_close(fc, h * s * torch.fft.fftn(s * x, dim=ax, norm="ortho"))
    ic = torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(x, dim=ax), dim=ax, norm="ortho"), dim=ax)
    torch.testing.assert_close(ic, h * s * torch.fft.ifftn(s * x, dim=ax, norm="ortho"))


def test_sign_helpers():
    assert even_grid((8, 6)) and not even_grid((8, 5))
    assert half_sign((8,)) == 1.0 and half_sign((6,)) == -1.0 and half_sign((6, 10)) == 1.0
    idx = torch.tensor([0, 1, 5, 6, 7, 23])
    # grid (4, 6): 5 -> (0, 5) odd, 6 -> (1, 0) odd, 7 -> (1, 1) even, 23 -> (3, 5) even
    np.testing.assert_array_equal(cell_signs(idx, (4, 6)).numpy(), [1, -1, -1, -1, 1, 1])
