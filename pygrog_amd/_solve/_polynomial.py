"""L2-optimal polynomial preconditioner (reference
src/pygrog/_solve/_polynomial.py; Johnson, Micchelli & Paul, SIAM J. Numer.
Anal. 20(2), 1983).

``p`` of degree d minimises ``int_l^L (1 - x p(x))^2 dx``.  With
``p(x) = sum_k c_k x^k`` the normal equations are ``M c = v`` with
``M_ij = int x^(i+j+2) dx`` and ``v_i = int x^(i+1) dx`` over [l, L], solved
here in closed form (the reference derives the same system with sympy).
"""

from __future__ import annotations

__all__ = ["PolynomialPreconditioner"]

import numpy as np
import torch


def _l2_optimal_coeffs(degree: int, lo: float, hi: float) -> list[float]:
    n = degree + 1

    def mom(p):  # int_lo^hi x^p dx
        return (hi ** (p + 1) - lo ** (p + 1)) / (p + 1)

    M = np.array([[mom(i + j + 2) for j in range(n)] for i in range(n)], dtype=np.float64)
    v = np.array([mom(i + 1) for i in range(n)], dtype=np.float64)
    return [float(c) for c in np.linalg.pinv(M) @ v]


class PolynomialPreconditioner:
    """``P(T) ~ T^-1`` on the spectrum bracket ``[l, L]`` of ``T = A^H A``."""

    def __init__(self, op, *, degree: int = 3, spectrum=None, n_power_iter: int = 10, sample_shape=None):
        if degree < 0:
            raise ValueError("degree must be >= 0")
        self._op = op
        self._sample_shape = tuple(int(s) for s in sample_shape) if sample_shape is not None else None
        normal_fn = getattr(op, "normal", None)
        if normal_fn is None:
            def normal_fn(x):
                return op.adjoint(op.forward(x))
        self._T = normal_fn
        if spectrum is None:
            lo, hi = 0.0, self._estimate_largest_eigenvalue(n_power_iter)
        else:
            lo, hi = float(spectrum[0]), float(spectrum[1])
        self.spectrum = (lo, hi)
        self.degree = int(degree)
        self.coeffs = _l2_optimal_coeffs(self.degree, lo, hi)

    def _estimate_largest_eigenvalue(self, n_iter: int) -> float:
        if self._sample_shape is not None:
            full = self._sample_shape
        else:
            image_shape = getattr(self._op, "image_shape", None) or getattr(self._op, "natural_shape", None)
            if image_shape is None:
                raise RuntimeError("Cannot estimate spectrum: pass `sample_shape=...` explicitly when `op` has no `image_shape`.")
            full = tuple(image_shape)
            if getattr(self._op, "smaps", None) is None:
                n_coils = getattr(self._op, "n_coils", None)
                if n_coils is None:
                    raise RuntimeError(
                        "Cannot infer coil dimension for spectrum estimate: pass `sample_shape=(n_coils, *image_shape)` explicitly."
                    )
                full = (int(n_coils), *full)
        device = getattr(self._op, "device", None) or torch.device("cpu")
        x = torch.randn(*full, dtype=torch.complex64, device=device)
        x = x / (torch.linalg.vector_norm(x) + 1e-30)
        eig = 0.0
        for _ in range(n_iter):
            y = self._T(x)
            eig = float(torch.linalg.vector_norm(y).item())
            x = y / (eig + 1e-30)
        return 1.05 * eig

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """Horner: ``c_0 x + T(c_1 x + T(... + T(c_d x)))``."""
        c = self.coeffs
        if self.degree == 0:
            return c[0] * x
        out = c[-1] * x
        for k in range(self.degree - 1, -1, -1):
            out = self._T(out) + c[k] * x
        return out

    def __call__(self, x):
        return self.apply(x)

    def __repr__(self):
        return f"PolynomialPreconditioner(degree={self.degree}, spectrum=({self.spectrum[0]:.3g}, {self.spectrum[1]:.3g}))"
