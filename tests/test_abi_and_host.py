"""CPU tests: the C ABI library loads and exports every symbol the header
declares; host-side logic (argument checks, shapes, solver maths on a dense
CPU operator, preconditioner coefficients, error messages)."""

import ctypes
import pathlib
import re

import numpy as np
import pytest
import torch

from pygrog_amd import _lib

HEADER = pathlib.Path(__file__).resolve().parents[1] / "include" / "pygrog_hip.h"


def _header_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(str(_lib.lib_path()))
    syms = _header_symbols()
    assert len(syms) >= 35
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, f"not exported: {missing}"
    # every declared entry point has a ctypes signature in the loader
    assert set(syms) <= set(_lib.exported_symbols())


def test_struct_layouts_match_header():
    # pg_frame: the ctypes mirror names the header's fields in the header's order
    hdr = HEADER.read_text()
    body = hdr[hdr.index("typedef struct pg_frame {"):hdr.index("} pg_frame;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"(\w+);", body)
    assert names == [f[0] for f in _lib.Frame._fields_]
    n_ptr = sum(1 for f in _lib.Frame._fields_ if f[1] is ctypes.c_void_p)
    assert ctypes.sizeof(_lib.Frame) == n_ptr * 8 + (len(names) - n_ptr) * 8
    # pg_plan_geom: 2 int32 + 6 x 3 fields + 128 x 3 int32
    assert ctypes.sizeof(_lib.PlanGeom) == 8 + 24 + 24 + 12 * 3 + 128 * 12 + 4  # trailing pad to 8


def test_library_version_and_errors():
    lib = _lib.load()
    assert lib.pg_abi_version() == 1
    # a host-side argument error is reported through pg_last_error (no GPU work)
    rc = lib.pg_grid_fp32(None, None, 0, 0, 1, 128, 0, 0, 0, None)
    assert rc != 0 and b"frame" in lib.pg_last_error()
    # 24-bit sample rows must start on 16-byte boundaries (checked before any
    # launch; the pointers are never dereferenced)
    assert lib.pg_transpose_cn_s24(4096, 32, 100, 4104, None) != 0 and b"16-byte" in lib.pg_last_error()
    assert lib.pg_transpose_frames_s24(4096, 4096, 1, 100, 32, 4104, None) != 0
    assert lib.pg_apply_grouped_bf16_s24(4104, 4096, 4096, 4096, 4096, 4096, 4096, 5, 10, 32, 3, 4096, None) != 0
    assert b"16-byte" in lib.pg_last_error()
    # the Gram products need 2-3 calibration axes of at least 2 points
    assert lib.pg_grappa_gram_f64(4096, 8, 3, _lib.dims([8, 1, 8]), 4096, None) != 0
    assert b"2 or more" in lib.pg_last_error()
    assert lib.pg_grappa_gram_f64(4096, 8, 1, _lib.dims([8]), 4096, None) != 0
    # the batched gather / apply / scatter entries reject bad sizes before any
    # launch; empty work returns 0 without touching the pointers
    assert lib.pg_gather_perm_stack_c64(4096, 16, 4096, 4096, 4096, 10, 2, 0, 1, 4096, None) != 0
    assert b"gather_perm_stack" in lib.pg_last_error()
    assert lib.pg_gather_perm_stack_c64(4096, 16, 4096, 4096, 4096, 0, 2, 3, 1, 4096, None) == 0
    assert lib.pg_sparse_apply_stack_c64(4096, 4096, 4096, 4096, 1, 2, 3, 10, 0, 5, None) != 0
    assert b"sparse_apply_stack" in lib.pg_last_error()
    assert lib.pg_sparse_apply_stack_c64(4096, 4096, 4096, 4096, 0, 2, 3, 10, 1, 5, None) == 0
    assert lib.pg_scatter_add_rows_c64(4096, 16, 4096, 4096, 4096, 10, 70000, None) != 0
    assert b"scatter_add_rows" in lib.pg_last_error()
    assert lib.pg_scatter_add_rows_c64(4096, 16, 4096, 4096, 4096, 10, 0, None) == 0
    assert lib.pg_grid_tile_cells(65) == 0
    assert _lib.tile_options(65) == [] and _lib.tile_options(0) == []
    # the table assembly takes up to 32 coils and 2 or 3 axes
    assert lib.pg_table_compose_f64(4096, 3, 11, 33, 0, 4096, None, None) != 0
    assert b"32 coils" in lib.pg_last_error()
    assert lib.pg_table_compose_f64(4096, 4, 11, 8, 0, 4096, None, None) != 0
    # the deferred grouping checks its sizes and pointers before any launch
    # (its host length pointer may be NULL; the synchronising call's may not)
    assert lib.pg_group_by_entry_async(4096, 4096, 4096, 0, 5, 128, 4096, 4096, 4096, 4096, 4096, 10, 4096, 1 << 20,
                                       None, None) != 0
    assert b"bad sizes" in lib.pg_last_error()
    assert lib.pg_group_by_entry(4096, 4096, 4096, 100, 5, 128, 4096, 4096, 4096, 4096, 4096, 10, 4096, 1 << 20,
                                 None, None) != 0
    for c, gs in ((3, 4), (8, 8), (16, 16), (32, 32), (64, 64)):
        tc = lib.pg_grid_tile_cells(c)
        opts = _lib.tile_options(c)
        assert tc in opts and opts == sorted(opts, reverse=True)
        # 256 is the wave-tile kernel (one wavefront per 64 cells, every coil
        # count); in the other shapes each thread of the 256-thread tile owns
        # one cell and at most 16 coils
        assert opts[0] == 256
        assert all(256 % t == 0 and 1 <= gs * t // 256 <= 16 for t in opts[1:])
    # a tile size that is not compiled is rejected before any launch
    rc = lib.pg_grid_fp32(1, 1, 1, 1, 16, 100, 0, 0, 0, None)
    assert rc != 0 and b"tile_cells" in lib.pg_last_error()
    # compact complex bf16 table: (re, im) per (co, ci), padded to 8 MT x 16 KS
    assert lib.pg_table_bf16_elems(10, 32, 2) == 10 * 2 * 32 * 32 * 2
    assert lib.pg_table_bf16_elems(3, 5, 1) == 3 * 1 * 8 * 16 * 2
    # FP32 packed table: coil pairs padded to half the coil group (4/8/16/32/64)
    assert lib.pg_table_packed_elems(10, 16) == 10 * 8 * 16
    assert lib.pg_table_packed_elems(3, 5) == 3 * 4 * 5
    assert lib.pg_table_packed_elems(2, 33) == 2 * 32 * 33
    assert lib.pg_table_packed_elems(1, 65) == 0 and lib.pg_table_packed_elems(0, 8) == 0
    # 24-bit sample rows: 2C values of 3 bytes, padded to whole 16-byte units
    assert [lib.pg_s24_row_bytes(c) for c in (0, 1, 2, 3, 8, 16, 20, 32, 64)] == [0, 16, 16, 32, 48, 96, 128, 192, 384]
    assert lib.pg_s24_row_bytes(-1) == -1
    # the BF16 modes use them where a row touches no more 128-byte lines than
    # a complex64 row: 32 and 20 coils yes, 16 and 8 coils no, never fp32
    from pygrog_amd.calib._grog import uses_s24

    assert uses_s24("bf16x3", 32) and uses_s24("bf16", 20)
    assert not uses_s24("bf16x3", 16) and not uses_s24("bf16x3", 8) and not uses_s24("fp32", 32)
    assert not uses_s24("bf16x3", 33) and not uses_s24("bf16x3c", 32)
    # block-floating-point rows: 16 bytes per 4 coils, 16-byte aligned buffers
    from pygrog_amd.calib._grog import row_format, sample_row_bytes

    assert [lib.pg_bfp_row_bytes(c) for c in (0, 1, 4, 5, 16, 32, 33)] == [0, 16, 16, 32, 64, 128, 144]
    assert lib.pg_bfp_row_bytes(-1) == -1
    assert [row_format("bf16x3c", c) for c in (1, 16, 32, 64)] == ["bfp"] * 4
    assert row_format("bf16x3", 32) == "s24" and row_format("fp32", 32) == "c64"
    assert sample_row_bytes(32, "bfp") == 128 and sample_row_bytes(32, "c64") == 256
    assert lib.pg_transpose_cn_bfp(4096, 32, 100, 4104, None) != 0 and b"16-byte" in lib.pg_last_error()
    assert lib.pg_transpose_frames_bfp(4096, 4096, 1, 100, 32, 4104, None) != 0
    assert lib.pg_apply_grouped_bf16_bfp(4104, 4096, 4096, 4096, 4096, 4096, 4096, 5, 10, 32, 3, 4096, None) != 0
    assert b"16-byte" in lib.pg_last_error()
    # block-floating-point V rows are reduced by the wave tiles only
    assert lib.pg_grid_reduce_v_bfp(4096, 4096, 1, 1, 32, 64, 0, 0, 0, None) != 0
    assert b"wave tiles" in lib.pg_last_error()
    from pygrog_amd.calib._grog import mode_tile_options, reduce_entry, v_format

    assert v_format("bf16x3c") == "bfp" and v_format("bf16x3") == "c64"
    assert reduce_entry("bf16x3c") == "pg_grid_reduce_v_bfp" and reduce_entry("bf16") == "pg_grid_reduce_v"
    assert mode_tile_options("bf16x3c", 32) == [256] and len(mode_tile_options("bf16x3", 32)) == 3


def test_kernel_offsets_and_quant_steps():
    from pygrog_amd.calib._grappa import quant_steps
    from pygrog_amd.calib._grog import _kernel_offsets

    assert _kernel_offsets(2, 2, "circle").shape == (3, 2)
    assert _kernel_offsets(3, 3, "sphere").shape == (19, 3)
    assert _kernel_offsets(2, 3, "cross").shape == (5, 2)
    with pytest.raises(ValueError, match="kernel_shape"):
        _kernel_offsets(2, 2, "hexagon")
    assert quant_steps(0.5, 1) == (10.0, 0.1, 11)
    assert quant_steps(0.5, None, 32) == (32.0, 1 / 32, 33)


def test_product_offsets_match_oracle():
    from oracle.plan import kernel_offsets as oko
    from pygrog_amd.calib._grog import _kernel_offsets

    for nd, w, s in [(2, 2, "circle"), (2, 3, "square"), (3, 2, "sphere"), (3, 3, "cross"), (3, 4, "cube")]:
        assert np.array_equal(_kernel_offsets(nd, w, s), oko(nd, w, s))


def test_maskedfft_requires_plan_or_explicit_args():
    from pygrog_amd.operator import MaskedFFT

    with pytest.raises(ValueError, match=r"plan.*required"):
        MaskedFFT()


def test_sparsefft_requires_plan_or_shapes():
    from pygrog_amd.operator import SparseFFT

    with pytest.raises(ValueError, match="grid_shape"):
        SparseFFT(indices=[0, 1], weights=[1, 1])


def test_no_gpu_fails_loudly():
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from pygrog_amd.calib import GrogInterpolator

    with pytest.raises(RuntimeError, match="HIP"):
        GrogInterpolator((8, 8), np.zeros((2, 4, 2), np.float32))


def test_stack_rejects_mismatch():
    from pygrog_amd.calib import GrogPlan

    with pytest.raises(ValueError, match="empty"):
        GrogPlan._stack([])
    a = GrogPlan(shape=(8, 8), oversamp=(1, 1), grid_shape=(8, 8), kernel_width=2, kernel_shape="circle", radius=0.5,
                 image_shape=(8, 8), grid_size=64, n_samples=4, natural_shape=(4,))
    b = GrogPlan(shape=(8, 8), oversamp=(1, 1), grid_shape=(9, 9), kernel_width=2, kernel_shape="circle", radius=0.5,
                 image_shape=(8, 8), grid_size=81, n_samples=4, natural_shape=(4,))
    with pytest.raises(ValueError, match="grid_shape"):
        GrogPlan._stack([a, b])


class _DenseOp:
    """y = A x on the CPU: the solvers only need forward/adjoint/normal."""

    def __init__(self, A):
        self.A = A
        self.toeplitz = True

    def forward(self, x):
        return self.A @ x

    def adjoint(self, y):
        return self.A.conj().T @ y

    def normal(self, x):
        return self.adjoint(self.forward(x))


def _problem(m=40, n=12, seed=0):
    g = torch.Generator().manual_seed(seed)
    A = torch.randn(m, n, dtype=torch.complex128, generator=g)
    b = torch.randn(m, dtype=torch.complex128, generator=g)
    return A, b


def test_cg_converges_to_normal_equation_solution():
    from pygrog_amd import cg

    A, b = _problem()
    damp = 0.1
    x = cg(_DenseOp(A), b, damp=damp, max_iter=60, tol=1e-14)
    ref = torch.linalg.solve(A.conj().T @ A + damp * torch.eye(12, dtype=A.dtype), A.conj().T @ b)
    torch.testing.assert_close(x, ref, rtol=1e-8, atol=1e-10)


def test_lsmr_converges_to_least_squares():
    from pygrog_amd import lsmr

    A, b = _problem(seed=1)
    x = lsmr(_DenseOp(A), b, max_iter=200, atol=1e-15, btol=1e-15)
    ref = torch.linalg.lstsq(A, b.unsqueeze(-1)).solution.squeeze(-1)
    torch.testing.assert_close(x, ref, rtol=1e-7, atol=1e-9)
    # damped: min ||Ax-b||^2 + damp^2 ||x||^2
    damp = 0.7
    x = lsmr(_DenseOp(A), b, damp=damp, max_iter=200, atol=1e-15, btol=1e-15)
    ref = torch.linalg.solve(A.conj().T @ A + damp**2 * torch.eye(12, dtype=A.dtype), A.conj().T @ b)
    torch.testing.assert_close(x, ref, rtol=1e-7, atol=1e-9)


def test_solve_dispatch_and_errors():
    from pygrog_amd import solve
    from pygrog_amd._solve import SolveMixin

    A, b = _problem(seed=2)
    op = _DenseOp(A)
    with pytest.raises(ValueError, match="Unknown solver method"):
        solve(op, b, method="bogus")

    class Op(SolveMixin, _DenseOp):
        pass

    o = Op(A)
    with pytest.raises(ValueError, match="Unknown method"):
        o.solve(b, method="bogus")
    x1 = o.solve(b, max_iter=5)
    x2 = o.solve(b, method="cg", max_iter=5)
    assert torch.equal(x1, x2)
    xn = o.solve(b.numpy(), method="cg", max_iter=5)
    assert isinstance(xn, np.ndarray)


def test_polynomial_coeffs_match_symbolic_l2_fit():
    sympy = pytest.importorskip("sympy")
    from pygrog_amd._solve._polynomial import _l2_optimal_coeffs

    deg, lo, hi = 3, 0.0, 2.5
    c = sympy.symbols(f"c0:{deg + 1}")
    x = sympy.symbols("x")
    p = sum(c[k] * x**k for k in range(deg + 1))
    J = sympy.integrate((1 - x * p) ** 2, (x, lo, hi))
    sol = sympy.solve([sympy.diff(J, ck) for ck in c], c)
    ref = [float(sol[ck]) for ck in c]
    np.testing.assert_allclose(_l2_optimal_coeffs(deg, lo, hi), ref, rtol=1e-8)


def test_polynomial_preconditioner_applies_horner():
    from pygrog_amd import PolynomialPreconditioner

    A, _ = _problem(seed=3)
    op = _DenseOp(A)
    op.image_shape = (12,)
    op.smaps = torch.ones(1)
    pc = PolynomialPreconditioner(op, degree=2, spectrum=(0.0, 100.0))
    x = torch.randn(12, dtype=torch.complex128)
    T = A.conj().T @ A
    c = pc.coeffs
    ref = c[0] * x + c[1] * (T @ x) + c[2] * (T @ (T @ x))
    torch.testing.assert_close(pc.apply(x), ref, rtol=1e-10, atol=1e-10)
    with pytest.raises(ValueError):
        PolynomialPreconditioner(op, degree=-1)


def test_lanczos_spectrum_bound():
    """The spectrum bound of the preconditioner (Lanczos, float64 inner
    products): never below the largest eigenvalue of a Hermitian T, equal
    to it once the Krylov space is exhausted; the coefficients then come
    from that bracket."""
    from pygrog_amd import PolynomialPreconditioner
    from pygrog_amd._solve._polynomial import _lanczos_upper_bound

    rng = np.random.default_rng(8)
    n = 40
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    lam = np.sort(rng.uniform(0.1, 5.0, n))
    T = torch.from_numpy(Q @ np.diag(lam) @ Q.conj().T)
    x0 = torch.from_numpy(rng.standard_normal(n) + 1j * rng.standard_normal(n))
    bounds = [_lanczos_upper_bound(lambda v: T @ v, x0, k) for k in (2, 5, 10, n)]
    assert all(b >= lam[-1] * (1 - 1e-12) for b in bounds)
    assert abs(bounds[-1] - lam[-1]) <= 1e-8 * lam[-1]
    assert _lanczos_upper_bound(lambda v: 0 * v, x0, 3) == 0.0

    class Op:
        image_shape = (n,)
        smaps = torch.ones(1)

        def normal(self, v):
            return T.to(v.dtype) @ v

    pc = PolynomialPreconditioner(Op(), degree=3, n_power_iter=n)
    assert pc.spectrum[0] == 0.0 and abs(pc.spectrum[1] - lam[-1]) <= 1e-5 * lam[-1]
    from pygrog_amd._solve._polynomial import _l2_optimal_coeffs

    assert pc.coeffs == _l2_optimal_coeffs(3, 0.0, pc.spectrum[1])


def test_resize_and_centered_fft_roundtrip():
    from pygrog_amd._base import fft, ifft
    from pygrog_amd._utils import resize

    x = torch.randn(6, 10, dtype=torch.complex64)
    torch.testing.assert_close(ifft(fft(x)), x, rtol=1e-5, atol=1e-5)
    big = resize(x, (8, 12))
    assert big.shape == (8, 12) and torch.equal(resize(big, (6, 10)), x)
    # centred FFT of a centred delta is flat with ortho norm
    d = torch.zeros(8, 8, dtype=torch.complex64)
    d[4, 4] = 1
    torch.testing.assert_close(fft(d), torch.full((8, 8), 1 / 8, dtype=torch.complex64))
