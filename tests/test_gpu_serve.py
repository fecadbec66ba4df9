"""Serving loop (GrogServer) on the HIP device: a Poisson stream of mixed 2D
and 3D jobs, each with its own table, admitted by HBM budget and gridded
in shared launches; every job's result equals the same job gridded alone,
bitwise, in every mode."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _jobs(device, n=10, seed=0, rate=400.0):
    from pygrog_amd import synth
    from pygrog_amd.serve import GridJob, poisson_arrivals

    rng = np.random.default_rng(seed)
    arr = poisson_arrivals(n, rate, seed)
    jobs = []
    for i in range(n):
        if i % 3 == 2:
            shape, C = (16, 16, 16), 4
            coords = synth.kooshball_3d(int(rng.integers(60, 160)), 32, first=int(rng.integers(0, 1000)))
            calib = synth.calibration(shape, C, (8, 8, 8), seed=seed + i, device=device)
            kw = dict(kernel_width=2, kernel_shape="sphere", oversamp=(1, 1, 1), precision=1)
        else:
            shape, C = (40, 40), 6
            coords = synth.radial_2d(int(rng.integers(6, 30)), 64, offset=float(rng.uniform(0, np.pi)))
            calib = synth.calibration(shape, C, (12, 12), seed=seed + i, device=device)
            kw = dict(kernel_width=3, quant=16)
        y = synth.complex_gaussian((C, coords.shape[0] * coords.shape[1]), seed=seed + 50 + i, device=device)
        jobs.append(GridJob(shape=shape, coords=coords, calib=calib, kspace=y, arrival=float(arr[i]), job_id=i,
                            lamda=1e-3, **kw))
    return jobs


def _alone(job, mode, device):
    from pygrog_amd.calib import GrogInterpolator

    g = GrogInterpolator(job.shape, job.coords, job.oversamp, job.kernel_width, job.kernel_shape, device=device)
    g.calc_interp_table(job.calib, lamda=job.lamda, precision=job.precision, quant=job.quant)
    return g.grid_device(job.kspace.reshape(1, 1, job.coils, -1).contiguous(), mode)


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "bf16"])
def test_serve_mix_equals_alone(device, mode):
    from pygrog_amd.serve import GrogServer

    jobs = _jobs(device, seed=1 + len(mode))
    with GrogServer(device, mode=mode, slots=3, keep_results=True) as srv:
        stats = srv.run(jobs, timeout=300)
    assert stats["jobs"] == len(jobs) and stats["p99_ms"] >= stats["p50_ms"] > 0
    assert stats["batches"] >= 1
    for j in jobs:
        assert j.t_admit >= j.arrival - 1e-6 and j.t_done >= j.t_ready >= j.t_admit
        ref = _alone(j, mode, device)
        assert torch.equal(j.result.reshape(-1), ref.reshape(-1)), f"job {j.job_id} differs from its run alone"


def test_serve_budget_queues_and_completes(device):
    from pygrog_amd.serve import GrogServer, estimate_job_bytes

    jobs = _jobs(device, n=8, seed=7, rate=2000.0)
    biggest = max(estimate_job_bytes(j, "fp32") for j in jobs)
    with GrogServer(device, mode="fp32", hbm_budget=int(1.5 * biggest), slots=4) as srv:
        stats = srv.run(jobs, timeout=300)
        assert stats["jobs"] == 8
        assert stats["peak_admitted_gb"] * 1e9 <= 1.5 * biggest + 1
        # the same server (and its worker threads) serves a second stream
        again = srv.run(_jobs(device, n=4, seed=8, rate=2000.0), timeout=300)
        assert again["jobs"] == 4
    with pytest.raises(ValueError, match="budget"):
        GrogServer(device, mode="fp32", hbm_budget=biggest // 4).run(_jobs(device, n=3, seed=9))
