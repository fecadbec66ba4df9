"""The torch custom-op layer on the CPU: the seven ops are registered under
``torch.ops.pygrog_amd`` with the reference's names (csrc/torch/module.cpp:
54-87), the in-place ones declare their mutated argument, the fake kernels
propagate shapes, and real CPU tensors are refused (no CPU kernel)."""

import pytest
import torch
from torch._subclasses.fake_tensor import FakeTensorMode

from pygrog_amd import _torch_ops


def test_registered_with_mutation_annotations():
    ns = torch.ops.pygrog_amd
    mutated = {"scatter_add": "grid", "scatter_add_binned": "grid", "psf_scatter_scalar": "psf",
               "psf_scatter_outer": "psf", "psf_scatter_outer_basis": "psf"}
    for name in _torch_ops.OP_NAMES:
        schema = getattr(ns, name).default._schema
        written = [a.name for a in schema.arguments if a.alias_info is not None and a.alias_info.is_write]
        assert written == ([mutated[name]] if name in mutated else []), (name, str(schema))
    assert set(dir(_torch_ops.ops())) == set(_torch_ops.OP_NAMES)
    with pytest.raises(AttributeError):
        _torch_ops.ops().not_an_op


def test_fake_kernels_propagate_shapes():
    with FakeTensorMode():
        d = torch.empty(10, 2, 4, dtype=torch.complex128)
        out = torch.ops.pygrog_amd.grog_interpolate(d, torch.empty(10, dtype=torch.int64),
                                                    torch.empty(3, 4, 4, dtype=torch.complex128))
        assert out.shape == (10, 2, 4) and out.dtype == torch.complex128
        g = torch.ops.pygrog_amd.gather(torch.empty(7, dtype=torch.complex64), torch.empty(5, dtype=torch.int64),
                                        torch.empty(5))
        assert g.shape == (5,) and g.dtype == torch.complex64
        grid = torch.empty(7, dtype=torch.complex64)
        assert torch.ops.pygrog_amd.scatter_add(grid, torch.empty(5, dtype=torch.complex64),
                                                torch.empty(5, dtype=torch.int64), torch.empty(5)) is None
        with pytest.raises(Exception, match="3-D"):
            torch.ops.pygrog_amd.grog_interpolate(torch.empty(10, 4, dtype=torch.complex64),
                                                  torch.empty(10, dtype=torch.int64),
                                                  torch.empty(3, 4, 4, dtype=torch.complex64))


def test_cpu_tensors_refused():
    with pytest.raises(RuntimeError, match="GPU only"):
        torch.ops.pygrog_amd.gather(torch.ones(4, dtype=torch.complex64), torch.tensor([0, 1]), torch.ones(2))
    with pytest.raises(RuntimeError, match="GPU only"):
        torch.ops.pygrog_amd.psf_scatter_scalar(torch.zeros(4), torch.tensor([0]), torch.ones(1))
