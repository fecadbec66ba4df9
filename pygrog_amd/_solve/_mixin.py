"""``op.solve`` mixin (reference src/pygrog/_solve/_mixin.py:84-158)."""

from __future__ import annotations

__all__ = ["SolveMixin"]

import numpy as np
import torch


def _is_cupy_array(x) -> bool:
    return x.__class__.__module__.startswith("cupy") and hasattr(x, "__dlpack__")


def _to_torch(x):
    if isinstance(x, torch.Tensor):
        return x
    if isinstance(x, np.ndarray):
        return torch.as_tensor(x)
    if _is_cupy_array(x):
        return torch.from_dlpack(x)
    return torch.as_tensor(x)


def _backend(x) -> str:
    if isinstance(x, torch.Tensor):
        return "torch"
    if isinstance(x, np.ndarray):
        return "numpy"
    if _is_cupy_array(x):
        return "cupy"
    return "torch"


def _from_torch(x: torch.Tensor, backend: str):
    if backend == "numpy":
        return x.detach().cpu().numpy()
    if backend == "cupy":
        import cupy as cp

        return cp.from_dlpack(x.detach())
    return x


class SolveMixin:
    """``solve(b, method=None, ...)``: CG when ``self.toeplitz`` (default),
    LSMR otherwise."""

    def solve(self, b, *, method: str | None = None, damp: float = 0.0, x0=None, max_iter: int = 20,
              tol: float = 1e-6, preconditioner=None, callback=None, **kwargs):
        from ._solvers import cg, lsmr

        backend = _backend(b)
        b_t = _to_torch(b)
        x0_t = None if x0 is None else _to_torch(x0)
        if method is None:
            method = "cg" if getattr(self, "toeplitz", False) else "lsmr"
        m = method.lower()
        if m == "cg":
            out = cg(self, b_t, damp=damp, x0=x0_t, max_iter=max_iter, tol=tol, preconditioner=preconditioner,
                     callback=callback, **kwargs)
            return _from_torch(out, backend)
        if m == "lsmr":
            out = lsmr(self, b_t, damp=damp, x0=x0_t, max_iter=max_iter, callback=callback, **kwargs)
            return _from_torch(out, backend)
        raise ValueError(f"Unknown method: {method!r}")
