"""Autograd wrappers around a SparseFFT (drop-in for the reference's
``grog_measure`` / ``grog_backproject``, src/pygrog/interop/_torch.py:66-128).

``grog_measure(x, op)`` applies A (image -> k-space, ``op.forward``) and
back-propagates with A^H (``op.adjoint``); ``grog_backproject(y, op)``
applies A^H and back-propagates with A.  Both run the HIP operators, so a
model that puts them in its graph trains through the GPU gridding.
"""

from __future__ import annotations

__all__ = ["grog_backproject", "grog_measure"]

import torch
from torch.autograd import Function


class _GrogMeasureFn(Function):
    @staticmethod
    def forward(ctx, x: torch.Tensor, op) -> torch.Tensor:
        ctx.op = op
        return op.forward(x)

    @staticmethod
    def backward(ctx, grad: torch.Tensor):
        return ctx.op.adjoint(grad.contiguous()), None


class _GrogBackprojectFn(Function):
    @staticmethod
    def forward(ctx, y: torch.Tensor, op) -> torch.Tensor:
        ctx.op = op
        ctx.y_shape = y.shape
        return op.adjoint(y)

    @staticmethod
    def backward(ctx, grad: torch.Tensor):
        # A applied to the image gradient, returned in the k-space's shape
        return ctx.op._adjoint_flat(grad.contiguous()).reshape(ctx.y_shape), None


def grog_measure(x: torch.Tensor, op) -> torch.Tensor:
    """A x with the gradient A^H."""
    return _GrogMeasureFn.apply(x, op)


def grog_backproject(y: torch.Tensor, op) -> torch.Tensor:
    """A^H y with the gradient A."""
    return _GrogBackprojectFn.apply(y, op)
