"""Multi-frame batched gridding (GrogFrameBatch) on the HIP device: every
frame of a batch with uneven sample counts equals its own interpolator's
grid bitwise (all modes) and the oracle within rel-L2 1e-5 (FP32); the
HIP-graph replay equals the eager result on every replay; autotuning the
launch shape leaves the bits unchanged."""

import numpy as np
import pytest
import torch

from oracle import grid as ogrid
from oracle import plan as oplan

pytestmark = pytest.mark.gpu

SHAPE, C, NREAD = (48, 48), 8, 96


def _batch(device, spokes=(5, 13, 2, 9, 1, 7), seed=0, coils=C):
    from pygrog_amd import synth
    from pygrog_amd.calib import GrogFrameBatch

    rng = np.random.default_rng(seed)
    coords = [synth.radial_2d(n, NREAD, offset=float(rng.uniform(0, np.pi))) for n in spokes]
    b = GrogFrameBatch(SHAPE, coords, kernel_width=3, device=device)
    calib = synth.calibration(SHAPE, coils, (16, 16), seed=seed + 1, device=device)
    b.calc_interp_table(calib, lamda=1e-3, quant=16)
    ys = [synth.complex_gaussian((coils, n * NREAD), seed=seed + 10 + f, device=device) for f, n in enumerate(spokes)]
    return b, coords, ys


@pytest.mark.parametrize("coils", [C, 20])
@pytest.mark.parametrize("mode", ["fp32", "bf16", "bf16x3"])
def test_batch_frames_equal_their_own_interpolator(device, mode, coils):
    """(20 coils: the BF16 modes go through 24-bit sample rows.)"""
    from pygrog_amd.calib._grog import uses_s24

    assert uses_s24(mode, coils) == (mode != "fp32" and coils == 20)
    b, coords, ys = _batch(device, coils=coils)
    y = b.pack(ys)
    out = b.grid(y, mode)
    torch.cuda.synchronize()
    for f, fr in enumerate(b.frames):
        alone = fr.grid_device(ys[f].reshape(1, 1, coils, -1).contiguous(), mode)
        assert torch.equal(out[f].reshape(coils, -1), alone.reshape(coils, -1)), f"frame {f} differs from its own grid"


def test_batch_fp32_vs_oracle(device):
    b, coords, ys = _batch(device, seed=3)
    out = b.grid(b.pack(ys), "fp32")
    for f, fr in enumerate(b.frames):
        p = oplan.attach_fft_plan(oplan.create_plan_single(SHAPE, coords[f], None, 3, "circle"))
        e = fr._interp_idx.reshape(p["distances"].shape[:-1]).cpu().numpy()
        ref, dens = ogrid.grid_reference(ys[f].cpu().numpy(), p, e, fr._interp_kernel.cpu().numpy())
        assert ogrid.rel_l2(out[f].reshape(C, -1).cpu().numpy(), ref) <= 1e-5
        np.testing.assert_allclose(b.density(f).reshape(-1).cpu().numpy(), dens, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("mode,coils", [("fp32", C), ("bf16x3", C), ("bf16x3", 20), ("bf16x3c", C)])
def test_batch_graph_replays_match_eager(device, mode, coils):
    b, _, ys = _batch(device, seed=5, coils=coils)
    y = b.pack(ys)
    eager = b.grid(y, mode).clone()
    graph, out = b.capture(y, mode)
    for k in range(3):
        out.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"replay {k} differs from the eager result"
        # eager work between replays, on the same buffers
        e2 = b.grid(y, mode)
        torch.cuda.synchronize()
        assert torch.equal(e2, eager)
    # new data in the captured input buffer is picked up by the replay
    y2 = b.pack([t * 2 for t in ys])
    y.copy_(y2)
    graph.replay()
    torch.cuda.synchronize()
    assert torch.equal(out, b.grid(y2, mode))


def test_batch_autotune_keeps_bits(device):
    b, _, ys = _batch(device, seed=7)
    y = b.pack(ys)
    ref = {m: b.grid(y, m).clone() for m in ("fp32", "bf16")}
    for m in ("fp32", "bf16"):
        res = b.autotune(y, m, reps=2)
        assert res["n_blocks"] >= 8
        assert torch.equal(b.grid(y, m), ref[m])


def test_batch_input_checks(device):
    b, _, ys = _batch(device, spokes=(3, 4), seed=9)
    with pytest.raises(ValueError):
        b.pack(ys[:1])
    with pytest.raises(ValueError):
        b.grid(torch.zeros(5, dtype=torch.complex64, device=device))
    with pytest.raises(ValueError):
        b.grid(b.pack(ys), "fp16")
