"""The seven native ops as torch custom operators, namespace ``pygrog_amd``.

The reference registers ``grog_interpolate``, ``scatter_add``,
``scatter_add_binned``, ``gather``, ``psf_scatter_scalar``,
``psf_scatter_outer`` and ``psf_scatter_outer_basis`` in a pybind11 module
(csrc/torch/module.cpp:54-87).  Here each is a ``torch.library`` operator
whose schema marks the in-place argument as mutated (``Tensor(a!)``), so
``torch.ops.pygrog_amd.<name>`` calls are visible to torch.compile, fake
tensor propagation and functionalisation.  The real kernels validate shapes
and dtypes with the reference's messages (mismatched dtypes raise instead of
being undefined behaviour) and enqueue the C-ABI entry points of
``include/pygrog_hip.h`` on the current stream; the fake kernels only
propagate shapes.  There is no CPU kernel: a CPU tensor raises.

``ops()`` returns an object with the seven names as attributes, the drop-in
for the reference's ``_get_torch_ext()`` module.
"""

from __future__ import annotations

__all__ = ["NAMESPACE", "TorchOps", "ops"]

import torch

from . import _lib

NAMESPACE = "pygrog_amd"
_T = _lib.TorchExt


@torch.library.custom_op(f"{NAMESPACE}::grog_interpolate", mutates_args=())
def grog_interpolate(data: torch.Tensor, indexes: torch.Tensor, kernel: torch.Tensor) -> torch.Tensor:
    """out[n, b, :] = kernel[indexes[n]] @ data[n, b, :] (grog_interp.cpp:112-149)."""
    return _T.grog_interpolate(data, indexes, kernel)


@grog_interpolate.register_fake
def _(data, indexes, kernel):
    torch._check(data.dim() == 3, lambda: f"data must be 3-D (nsamples, nbatches, ncoils), got {data.dim()}-D")
    torch._check(kernel.dtype == data.dtype, lambda: f"kernel dtype {kernel.dtype} != data dtype {data.dtype}")
    return torch.empty_like(data, memory_format=torch.contiguous_format)


@torch.library.custom_op(f"{NAMESPACE}::scatter_add", mutates_args=("grid",))
def scatter_add(grid: torch.Tensor, data: torch.Tensor, indices: torch.Tensor, weights: torch.Tensor) -> None:
    """grid[indices[i]] += weights[i] * data[i], in place (sparse_ops.cpp:233-267)."""
    _T.scatter_add(grid, data, indices, weights)


@scatter_add.register_fake
def _(grid, data, indices, weights):
    return None


@torch.library.custom_op(f"{NAMESPACE}::scatter_add_binned", mutates_args=("grid",))
def scatter_add_binned(grid: torch.Tensor, data: torch.Tensor, indices: torch.Tensor, weights: torch.Tensor,
                       bin_starts: torch.Tensor, bin_size: int) -> None:
    """scatter_add with the reference's bin arguments (sparse_ops.cpp:299-340)."""
    _T.scatter_add_binned(grid, data, indices, weights, bin_starts, bin_size)


@scatter_add_binned.register_fake
def _(grid, data, indices, weights, bin_starts, bin_size):
    return None


@torch.library.custom_op(f"{NAMESPACE}::gather", mutates_args=())
def gather(grid: torch.Tensor, indices: torch.Tensor, weights: torch.Tensor) -> torch.Tensor:
    """out[i] = weights[i] * grid[indices[i]] (sparse_ops.cpp:269-297)."""
    return _T.gather(grid, indices, weights)


@gather.register_fake
def _(grid, indices, weights):
    torch._check(indices.dim() == 1, lambda: "indices must be 1-D")
    return grid.new_empty((indices.shape[0],))


@torch.library.custom_op(f"{NAMESPACE}::psf_scatter_scalar", mutates_args=("psf",))
def psf_scatter_scalar(psf: torch.Tensor, indices: torch.Tensor, w_sq: torch.Tensor) -> None:
    """psf[indices[i]] += w_sq[i], in place (toep_psf.cpp:242-274)."""
    _T.psf_scatter_scalar(psf, indices, w_sq)


@psf_scatter_scalar.register_fake
def _(psf, indices, w_sq):
    return None


@torch.library.custom_op(f"{NAMESPACE}::psf_scatter_outer", mutates_args=("psf",))
def psf_scatter_outer(psf: torch.Tensor, indices: torch.Tensor, w_sq: torch.Tensor,
                      basis_per_sample: torch.Tensor) -> None:
    """psf[idx, l, l'] += w_sq * conj(b[l]) * b[l'], in place (toep_psf.cpp:276-320)."""
    _T.psf_scatter_outer(psf, indices, w_sq, basis_per_sample)


@psf_scatter_outer.register_fake
def _(psf, indices, w_sq, basis_per_sample):
    return None


@torch.library.custom_op(f"{NAMESPACE}::psf_scatter_outer_basis", mutates_args=("psf",))
def psf_scatter_outer_basis(psf: torch.Tensor, indices: torch.Tensor, w_sq: torch.Tensor, basis: torch.Tensor,
                            time_index: torch.Tensor) -> None:
    """psf[idx, k, k'] += w_sq * conj(B[k, t]) * B[k', t], in place (toep_psf.cpp:322-367)."""
    _T.psf_scatter_outer_basis(psf, indices, w_sq, basis, time_index)


@psf_scatter_outer_basis.register_fake
def _(psf, indices, w_sq, basis, time_index):
    return None


OP_NAMES = ("grog_interpolate", "scatter_add", "scatter_add_binned", "gather", "psf_scatter_scalar",
            "psf_scatter_outer", "psf_scatter_outer_basis")


class TorchOps:
    """``torch.ops.pygrog_amd`` under the names of the reference's
    ``_pygrog_torch`` module."""

    def __getattr__(self, name):
        if name in OP_NAMES:
            return getattr(getattr(torch.ops, NAMESPACE), name)
        raise AttributeError(name)

    def __dir__(self):
        return list(OP_NAMES)


_ops = TorchOps()


def ops() -> TorchOps:
    return _ops
