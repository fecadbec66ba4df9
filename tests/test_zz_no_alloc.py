"""The allocation-free inner product (pg_vdot_c64_f64_noalloc) against the
workspace entry point it shares its kernels with: the same FP64 bits, on
two calls in a row (the second on a smaller input, so that the scratch
array still holds the first call's partials), through the C ABI, and the
size and argument errors."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _vec(n, seed, device):
    g = torch.Generator(device=device).manual_seed(seed)
    return torch.complex(torch.randn(n, generator=g, device=device),
                         torch.randn(n, generator=g, device=device)).contiguous()


def test_noalloc_vdot_equals_workspace_entry(device):
    from pygrog_amd import _lib

    lib = _lib.load()
    st = _lib.stream(device)
    ws = torch.empty(int(lib.pg_vdot_workspace_bytes()), dtype=torch.uint8, device=device)
    for n, seed in ((3_000_001, 1), (70_001, 2), (5, 3), (0, 4)):
        a, b = _vec(max(n, 1), seed, device)[:n], _vec(max(n, 1), seed + 10, device)[:n]
        ref = torch.full((2,), -1.0, dtype=torch.float64, device=device)
        got = torch.full((2,), -2.0, dtype=torch.float64, device=device)
        _lib.call("pg_vdot_c64_f64", _lib.ptr(a), _lib.ptr(b), n, _lib.ptr(ws), _lib.ptr(ref), st)
        _lib.call("pg_vdot_c64_f64_noalloc", _lib.ptr(a), _lib.ptr(b), n, _lib.ptr(got), st)
        torch.cuda.synchronize(device)
        assert torch.equal(got, ref), (n, got, ref)
        if n:
            exact = np.vdot(a.cpu().numpy().astype(np.complex128), b.cpu().numpy().astype(np.complex128))
            assert abs(complex(got[0].item(), got[1].item()) - exact) <= 1e-9 * max(1.0, abs(exact))
        else:
            assert got.tolist() == [0.0, 0.0]


def test_noalloc_vdot_rejects_bad_sizes(device):
    from pygrog_amd import _lib

    lib = _lib.load()
    out = torch.zeros(2, dtype=torch.float64, device=device)
    a = _vec(8, 5, device)
    assert lib.pg_vdot_c64_f64_noalloc(_lib.ptr(a), _lib.ptr(a), -1, _lib.ptr(out), _lib.stream(device)) != 0
    assert b"bad arguments" in lib.pg_last_error()
    assert lib.pg_vdot_c64_f64_noalloc(_lib.ptr(a), _lib.ptr(a), (1 << 40) + 1, _lib.ptr(out),
                                       _lib.stream(device)) != 0
    assert b"exceeds the maximum" in lib.pg_last_error()
