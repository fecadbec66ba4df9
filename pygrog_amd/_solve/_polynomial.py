"""L2-optimal polynomial preconditioner for the normal operator ``T = A^H A``.

Drop-in for the reference's ``PolynomialPreconditioner``
(src/pygrog/_solve/_polynomial.py:67-207): same constructor signature,
attributes (``degree``, ``spectrum``, ``coeffs`` in the power basis) and
error messages.  The computation is this package's own:

* coefficients -- ``p`` of degree d minimises ``int_l^L (1 - x p(x))^2 dx``
  (Johnson, Micchelli & Paul, SIAM J. Numer. Anal. 20(2), 1983); with
  ``p(x) = sum_k c_k x^k`` the normal equations ``M c = v`` have
  ``M_ij = int x^(i+j+2)`` and ``v_i = int x^(i+1)`` over [l, L], formed
  from closed-form moments in float64;
* upper spectrum bound -- a few Lanczos steps on ``T`` (fully
  reorthogonalised) with the inner products accumulated in float64 (on a
  HIP device by the library's FP64 reduction), giving the largest Ritz
  value ``theta`` of the tridiagonal matrix and the last off-diagonal
  ``beta``; ``theta + beta`` is the spectrum's upper end (the residual
  bound of the Ritz pair), exact once the Krylov space is exhausted;
* application -- ``P(T) x = sum_k c_k T^k x`` accumulated term by term as
  the Krylov powers ``T^k x`` are formed: ``degree`` applications of ``T``.
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


def _inner64(a: torch.Tensor, b: torch.Tensor) -> complex:
    """<a, b> (a conjugated) accumulated in float64."""
    if a.is_cuda and a.dtype == torch.complex64 and b.dtype == torch.complex64 and a.device == b.device:
        from ._solvers import _vdot_c64_f64

        r = _vdot_c64_f64(a, b).cpu()
        return complex(float(r[0]), float(r[1]))
    return complex(torch.vdot(a.reshape(-1).to(torch.complex128), b.reshape(-1).to(torch.complex128)).item())


def _lanczos_upper_bound(T, x0: torch.Tensor, steps: int) -> float:
    """theta_max + beta_last of ``steps`` Lanczos steps on the Hermitian T
    from x0 (each new vector orthogonalised against all previous ones, so
    the recurrence keeps its exact-arithmetic meaning); 0 for a zero
    operator."""
    nrm = abs(_inner64(x0, x0)) ** 0.5
    if nrm == 0.0:
        return 0.0
    basis = [x0 / nrm]
    alphas, betas = [], []
    for _ in range(max(int(steps), 1)):
        v = basis[-1]
        w = T(v)
        a = _inner64(v, w).real
        for _ in range(2):  # two Gram-Schmidt passes against the whole basis
            for q in basis:
                w = w - _inner64(q, w) * q
        beta = abs(_inner64(w, w)) ** 0.5
        alphas.append(a)
        betas.append(beta)
        if beta <= 1e-12 * max(abs(a), 1e-300):
            break
        basis.append(w / beta)
    k = len(alphas)
    J = np.diag(np.asarray(alphas)) + np.diag(np.asarray(betas[:k - 1]), 1) + np.diag(np.asarray(betas[:k - 1]), -1)
    theta = float(np.linalg.eigvalsh(J).max())
    return max(theta, 0.0) + betas[-1]


class PolynomialPreconditioner:
    """``P(T) ~ T^-1`` on the spectrum bracket ``[l, L]`` of ``T = A^H A``
    (reference signature: ``op``, ``degree``, ``spectrum``, ``n_power_iter``
    -- here the Lanczos steps of the bound -- and ``sample_shape``)."""

    def __init__(self, op, *, degree: int = 3, spectrum=None, n_power_iter: int = 10, sample_shape=None):
        if degree < 0:
            raise ValueError("degree must be >= 0")
        self._op = op
        self._sample_shape = tuple(int(s) for s in sample_shape) if sample_shape is not None else None
        normal = getattr(op, "normal", None)
        self._T = normal if normal is not None else (lambda x: op.adjoint(op.forward(x)))
        if spectrum is not None:
            lo, hi = float(spectrum[0]), float(spectrum[1])
        else:
            lo, hi = 0.0, self._estimate_largest_eigenvalue(n_power_iter)
        self.spectrum = (lo, hi)
        self.degree = int(degree)
        self.coeffs = _l2_optimal_coeffs(self.degree, lo, hi)

    def _probe_shape(self) -> tuple:
        """Shape of T's input: ``sample_shape``, else the operator's image
        (or natural) shape, with the coil axis in front without SENSE maps."""
        if self._sample_shape is not None:
            return self._sample_shape
        shape = getattr(self._op, "image_shape", None) or getattr(self._op, "natural_shape", None)
        if shape is None:
            raise RuntimeError("Cannot estimate spectrum: pass `sample_shape=...` explicitly when `op` has no "
                               "`image_shape`.")
        shape = tuple(int(s) for s in shape)
        if getattr(self._op, "smaps", None) is not None:
            return shape
        n_coils = getattr(self._op, "n_coils", None)
        if n_coils is None:
            raise RuntimeError("Cannot infer coil dimension for spectrum estimate: pass "
                               "`sample_shape=(n_coils, *image_shape)` explicitly.")
        return (int(n_coils), *shape)

    def _estimate_largest_eigenvalue(self, n_iter: int) -> float:
        shape = self._probe_shape()
        dev = getattr(self._op, "device", None)
        if dev is None or torch.device(dev).type != "cuda":
            dev = torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")
        gen = torch.Generator(device=dev).manual_seed(0x5EED)
        x0 = torch.complex(torch.randn(shape, generator=gen, device=dev),
                           torch.randn(shape, generator=gen, device=dev))
        return _lanczos_upper_bound(self._T, x0, n_iter)

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """``sum_k c_k T^k x``: the Krylov powers formed one after the other."""
        c = self.coeffs
        out = c[0] * x
        tk = x
        for ck in c[1:]:
            tk = self._T(tk)
            out = out + ck * tk
        return out

    def __call__(self, x):
        return self.apply(x)

    def __repr__(self):
        return f"PolynomialPreconditioner(degree={self.degree}, spectrum=({self.spectrum[0]:.3g}, {self.spectrum[1]:.3g}))"
