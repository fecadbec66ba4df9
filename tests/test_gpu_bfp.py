"""The block-floating-point rows of the bf16x3c mode on the HIP device: the
transposes' bytes against the oracle's encoder, the apply's V rows against
the oracle's encoding of the complex64-V path on the same values, the
reduction against pg_grid_reduce_v on the decoded V, frames and batches,
and the mode against the FP64 oracle."""

import numpy as np
import pytest
import torch

from oracle import grid as ogrid
from oracle import plan as oplan
from oracle import rows

pytestmark = pytest.mark.gpu


def _rows_of(y: np.ndarray) -> np.ndarray:
    C, N = y.shape
    return np.stack([y.real.T, y.imag.T], axis=-1).reshape(N, 2 * C).astype(np.float32)


@pytest.mark.parametrize("C,N", [(32, 4096), (16, 1000), (3, 77), (40, 130), (1, 65), (6, 64)])
def test_transpose_bfp_rows(device, C, N):
    """pg_transpose_cn_bfp (full 64 x 32 tiles and ragged ones) writes every
    group's 16 bytes as the oracle encodes them -- zeros, subnormals, the
    largest float, infinities, NaN, half-step ties and groups whose maximum
    rounds past 16383 included; pg_transpose_frames_bfp writes each frame's
    rows at row_bytes x its sample offset."""
    from pygrog_amd import _lib

    lib = _lib.load()
    rb = int(lib.pg_bfp_row_bytes(C))
    assert rb == 16 * ((C + 3) // 4)
    rng = np.random.default_rng(C * 13 + N)
    re = (rng.standard_normal((C, N)) * np.exp2(rng.integers(-40, 40, (1, N)))).astype(np.float32)
    im = (rng.standard_normal((C, N)) * np.exp2(rng.integers(-40, 40, (1, N)))).astype(np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1e-40, -3e-39, 3.4028235e38,
                        np.float32(16383.75 / 8192), np.float32(16383.25 / 8192), np.float32(2.0 ** -14),
                        np.float32(3 * 2.0 ** -14)], dtype=np.float32)
    re.reshape(-1)[:special.size] = special
    im.reshape(-1)[-special.size:] = special
    y = np.empty((C, N), dtype=np.complex64)
    y.real, y.imag = re, im
    want = rows.bfp15_encode(_rows_of(y))
    yd = torch.from_numpy(y).to(device)
    print(f"transpose_bfp C={C} N={N}", flush=True)
    yb = torch.zeros(N * rb // 2, dtype=torch.int16, device=device)
    _lib.call("pg_transpose_cn_bfp", _lib.ptr(yd), C, N, _lib.ptr(yb), _lib.stream(device))
    got = yb.cpu().numpy().view(np.uint16).reshape(N, rb // 2)
    assert np.array_equal(got, want)
    n0 = N // 3
    off = torch.tensor([0, n0, N], dtype=torch.int64, device=device)
    yf = torch.from_numpy(np.concatenate([y[:, :n0].reshape(-1), y[:, n0:].reshape(-1)])).to(device)
    ybf = torch.zeros(N * rb // 2, dtype=torch.int16, device=device)
    _lib.call("pg_transpose_frames_bfp", _lib.ptr(yf), _lib.ptr(off), 2, N - n0, C, _lib.ptr(ybf),
              _lib.stream(device))
    assert np.array_equal(ybf.cpu().numpy().view(np.uint16).reshape(N, rb // 2), want)


def _radial(n_sp, n_rd, seed):
    from pygrog_amd import synth

    return synth.radial_2d(n_sp, n_rd, offset=float(np.random.default_rng(seed).uniform(0, np.pi)))


@pytest.mark.parametrize("C", [6, 16, 32])
def test_bf16x3c_apply_and_reduce_exact(device, C):
    """bf16x3c end to end, step by step.  The apply: for inputs on the
    block-floating-point grid every row format holds the same values and the
    split x = hi + lo is exact, so each V row of pg_apply_grouped_bf16_bfp
    is the oracle's encoding of the complex64 V row bf16x3 computes from the
    same values (24-bit rows at 6 and 32 coils, complex64 rows at 16).  The
    reduction: pg_grid_reduce_v_bfp grids the bits pg_grid_reduce_v grids
    from the oracle's decoding of those V rows."""
    from pygrog_amd.calib import GrogInterpolator
    from pygrog_amd.calib._grog import row_format

    rng = np.random.default_rng(900 + C)
    shape = (64, 64)
    coords = _radial(40, 128, seed=C)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    calib = (rng.standard_normal((C, 16, 16)) + 1j * rng.standard_normal((C, 16, 16))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=1e-2, quant=16)
    comp = g.plan._compiled
    assert comp.hot[0]["n_hot"] > 0  # the hot pre-pass decodes V too
    n = coords.shape[0] * coords.shape[1]
    y = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    q = rows.bfp15_decode(rows.bfp15_encode(_rows_of(y)), C).reshape(n, C, 2)
    yq = np.empty((C, n), dtype=np.complex64)
    yq.real, yq.imag = q[..., 0].T, q[..., 1].T
    assert row_format("bf16x3c", C) == "bfp" and row_format("bf16x3", C) in ("s24", "c64")
    yy = torch.from_numpy(yq).to(device).reshape(1, 1, C, -1).contiguous()
    cc, c3 = {}, {}
    a = g.grid_device(yy, "bf16x3c", program_cache=cc).clone()
    g.grid_device(yy, "bf16x3", program_cache=c3)
    torch.cuda.synchronize()
    (vb, _, _), = cc.values()
    (v64, yt3, _), = c3.values()
    R, ng = comp.n_valid[0], (C + 3) // 4
    got = vb.cpu().numpy().view(np.uint16).reshape(-1, 8 * ng)[:R]
    vr = v64.cpu().numpy().reshape(-1, C)[:R]
    want = rows.bfp15_encode(np.stack([vr.real, vr.imag], axis=-1).reshape(R, 2 * C))
    assert np.array_equal(got, want)
    dec = rows.bfp15_decode(got, C).reshape(R, C, 2)
    vd = torch.zeros_like(v64)
    vd[:R * C] = torch.from_numpy((dec[..., 0] + 1j * dec[..., 1]).astype(np.complex64).reshape(-1)).to(device)
    out = torch.full_like(a, complex(float("nan"), 0.0))
    g.grid_program(yt3, out, 1, vbuf=vd, mode="bf16x3").launch("pg_grid_reduce_v")
    torch.cuda.synchronize()
    assert torch.equal(out, a)


def test_bf16x3c_vs_fp64_and_batch(device):
    """bf16x3c on a 2D radial case within 2e-4 of the FP64 oracle (input and
    V rows each within 2^-14 of their 4-coil group's largest magnitude:
    about 7e-5 on Gaussian data), and a frame batch (GrogFrameBatch, one
    launch per kernel) equal to each frame gridded alone, bitwise."""
    from pygrog_amd.calib import GrogFrameBatch, GrogInterpolator

    rng = np.random.default_rng(77)
    shape, C = (96, 96), 16
    coords = _radial(64, 192, seed=5)
    g = GrogInterpolator(shape, coords, kernel_width=3, device=device)
    calib = (rng.standard_normal((C, 16, 16)) + 1j * rng.standard_normal((C, 16, 16))).astype(np.complex64)
    g.calc_interp_table(calib, lamda=1e-2, quant=32, keep_fp64=True)
    p = oplan.attach_fft_plan(oplan.create_plan_single(shape, coords, None, 3, "circle"))
    e = g._interp_idx.reshape(p["distances"].shape[:-1]).cpu().numpy()
    n = coords.shape[0] * coords.shape[1]
    y = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    g64 = ogrid.grid_fp64(y, p, e, g.plan._compiled.table64.cpu().numpy())
    yy = torch.from_numpy(y).to(device).reshape(1, 1, C, -1).contiguous()
    rel = ogrid.rel_l2(g.grid_device(yy, "bf16x3c").reshape(C, -1).cpu().numpy(), g64)
    print(f"bf16x3c rel-L2 vs FP64 {rel:.2e}", flush=True)
    assert rel <= 2e-4

    frames = [_radial(int(k), 128, seed=40 + i) for i, k in enumerate(rng.integers(5, 20, 5))]
    fb = GrogFrameBatch(shape, frames, kernel_width=3, device=device)
    fb.calc_interp_table(calib, lamda=1e-2, quant=32)
    ys = [(rng.standard_normal((C, f.shape[0] * f.shape[1])) + 1j * rng.standard_normal((C, f.shape[0] * f.shape[1])))
          .astype(np.complex64) for f in frames]
    out = fb.grid(ys, "bf16x3c")
    for i, f in enumerate(frames):
        gi = GrogInterpolator(shape, f, kernel_width=3, device=device)
        gi.calc_interp_table(calib, lamda=1e-2, quant=32)
        alone = gi.grid_device(torch.from_numpy(ys[i]).to(device).reshape(1, 1, C, -1).contiguous(), "bf16x3c")
        assert torch.equal(out[i].reshape(-1), alone.reshape(-1)), i
