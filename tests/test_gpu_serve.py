"""Serving loop (GrogServer) on the HIP device: a Poisson stream of mixed 2D
and 3D jobs, each with its own table, admitted by HBM budget and gridded
in shared launches; every job's result equals the same job gridded alone,
bitwise, in every mode."""

import dataclasses

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


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "bf16", "bf16x3c"])
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
    # a budget below every job: each is rejected alone, the run still ends
    with GrogServer(device, mode="fp32", hbm_budget=biggest // 4) as srv:
        st = srv.run(_jobs(device, n=3, seed=9), timeout=60)
    assert st["jobs"] == 0 and len(st["rejected"]) == 3 and all("budget" in e for _, e in st["rejected"])


@pytest.mark.parametrize("mode", ["fp32", "bf16x3"])
def test_serve_failures_isolated(device, monkeypatch, mode):
    """Malformed requests (k-space of the wrong length, a stacked
    trajectory), an over-budget job, a preparation that raises and an
    injected allocation failure each fail their own job only; every other
    job's result equals its run alone, bitwise."""
    from pygrog_amd import synth
    from pygrog_amd.serve import GrogServer, estimate_job_bytes

    jobs = _jobs(device, n=10, seed=21, rate=1000.0)
    jobs[1].kspace = jobs[1].kspace[:, :-3].contiguous()
    jobs[4].coords = np.stack([jobs[4].coords, jobs[4].coords])
    jobs[6].job_id = 106      # its table fit raises
    jobs[8].job_id = 108      # its buffers fail to allocate
    budget = 3 * max(estimate_job_bytes(j, mode) for j in jobs if j.job_id not in (1, 4))
    big = synth.complex_gaussian((4, 300000), seed=3, device=device)
    jobs.append(dataclasses.replace(jobs[2], kspace=big, coords=synth.kooshball_3d(1000, 300), job_id=99,
                                    arrival=0.0))
    with GrogServer(device, mode=mode, hbm_budget=budget, slots=3, keep_results=True) as srv:
        real_alloc = srv._alloc
        current = {}

        def alloc(n, dtype):
            if current.get("job") == 108:
                raise RuntimeError("injected: hipMalloc out of memory")
            return real_alloc(n, dtype)

        real_stage = srv._stage_job

        def stage(job, st, groups, keep):
            current["job"] = job.job_id
            return real_stage(job, st, groups, keep)

        real_fit = srv._fit_finish

        def fit(job, launched):
            if job.job_id == 106:
                raise RuntimeError("injected: table fit failed")
            return real_fit(job, launched)

        monkeypatch.setattr(srv, "_fit_finish", fit)
        monkeypatch.setattr(srv, "_alloc", alloc)
        monkeypatch.setattr(srv, "_stage_job", stage)
        stats = srv.run(jobs, timeout=300)
    status = {j.job_id: j.status for j in jobs}
    assert {i for i, s in status.items() if s == "rejected"} == {1, 4, 99}, stats["rejected"]
    assert {i for i, s in status.items() if s == "failed"} == {106, 108}, stats["failed"]
    assert stats["jobs"] == 6
    for j in jobs:
        if j.status == "done":
            ref = _alone(j, mode, device)
            assert torch.equal(j.result.reshape(-1), ref.reshape(-1)), f"job {j.job_id} differs from its run alone"
