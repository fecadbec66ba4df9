"""Host-side pieces of the serving loop (no GPU): arrival process and the
admission estimate."""

import numpy as np
import torch

from pygrog_amd.serve import GridJob, estimate_job_bytes, poisson_arrivals


def test_poisson_arrivals_rate_and_order():
    t = poisson_arrivals(20000, 50.0, seed=3)
    assert np.all(np.diff(t) > 0)
    assert abs(len(t) / t[-1] - 50.0) < 1.5
    assert np.array_equal(t, poisson_arrivals(20000, 50.0, seed=3))


def _job(n_samples, C=8, shape=(64, 64), **kw):
    coords = np.zeros((n_samples, len(shape)), np.float32)
    return GridJob(shape=shape, coords=coords, calib=torch.zeros(C, 8, 8, dtype=torch.complex64),
                   kspace=torch.zeros(C, n_samples, dtype=torch.complex64), **kw)


def test_estimate_grows_with_work_and_mode():
    small, big = _job(1000), _job(100000)
    assert estimate_job_bytes(big, "fp32") > estimate_job_bytes(small, "fp32")
    assert estimate_job_bytes(big, "bf16x3") > estimate_job_bytes(big, "fp32")
    # the estimate covers at least the k-space, its sample-major copy and the grid
    C, N, G = 8, 100000, 64 * 64
    assert estimate_job_bytes(big, "fp32") >= 2 * 8 * C * N + 8 * C * G
    # a finer quantisation means a larger table
    assert estimate_job_bytes(_job(1000, quant=64), "fp32") > estimate_job_bytes(_job(1000, quant=8), "fp32")


# -- failure isolation of the scheduling loop (fake device work) ---------------
import concurrent.futures as cf  # noqa: E402
import time  # noqa: E402

import pytest  # noqa: E402

from pygrog_amd.serve import serve_loop, summarize, validate_job  # noqa: E402


def _radial_job(job_id, n_sp=4, n_rd=16, C=4, arrival=0.0, **kw):
    r = (np.arange(n_rd) - n_rd // 2) / n_rd
    th = np.arange(n_sp) * 2.0
    coords = np.stack([np.outer(np.cos(th), r), np.outer(np.sin(th), r)], -1).astype(np.float32)
    return GridJob(shape=(32, 32), coords=coords, calib=torch.zeros(C, 8, 8, dtype=torch.complex64),
                   kspace=torch.zeros(C, n_sp * n_rd, dtype=torch.complex64), job_id=job_id, arrival=arrival,
                   kernel_width=3, quant=16, **kw)


def test_validate_job_names_each_malformed_request():
    assert validate_job(_radial_job(0)) is None
    j = _radial_job(1)
    j.kspace = torch.zeros(4, 63, dtype=torch.complex64)
    assert "63 samples" in validate_job(j)
    j = _radial_job(2)
    j.coords = np.zeros((2, 4, 16, 2), np.float32)  # a stack of trajectories
    assert "stack" in validate_job(j)
    j = _radial_job(3)
    j.calib = torch.zeros(5, 8, 8, dtype=torch.complex64)
    assert "calibration has 5" in validate_job(j)
    j = _radial_job(4, C=65)
    assert "64 coils" in validate_job(j)
    j = _radial_job(5)
    j.kernel_shape = "hexagon"
    assert validate_job(j).startswith("kernel")
    j = _radial_job(8)
    j.kernel_width = 0
    assert "no neighbours" in validate_job(j)
    j = _radial_job(6)
    j.kspace = torch.zeros(4, 64)  # real
    assert "complex" in validate_job(j)
    j = _radial_job(7)
    j.coords = np.zeros((64, 3), np.float32)
    assert "2 axes" in validate_job(j)


class _Handle:
    def __init__(self, t_done):
        self.t_done = t_done

    def query(self):
        return time.perf_counter() >= self.t_done


class _FakeDevice:
    """Stands in for the GPU: preparation takes a few ms and may raise,
    every launch allocates each job's buffers through ``alloc`` (the fake
    allocator) and completes a few ms later."""

    def __init__(self, fail_prepare=(), fail_alloc=()):
        self.fail_prepare = set(fail_prepare)
        self.fail_alloc = set(fail_alloc)
        self.allocated = 0
        self.launched = []

    def alloc(self, job):
        if job.job_id in self.fail_alloc:
            raise MemoryError(f"injected: out of device memory for job {job.job_id}")
        self.allocated += 1

    def prepare(self, job, slot):
        time.sleep(0.002)
        if job.job_id in self.fail_prepare:
            raise RuntimeError("plan compiler: injected failure")
        return job

    def launch(self, batch):
        bad = {}
        for job in batch:
            try:
                self.alloc(job)
            except MemoryError as exc:
                bad[job] = f"launch failed: MemoryError: {exc}"
        self.launched.append([j.job_id for j in batch if j not in bad])
        return _Handle(time.perf_counter() + 0.003), None, bad


def _run(jobs, dev, budget=10**9, estimate=lambda j: 1000):
    with cf.ThreadPoolExecutor(max_workers=3) as pool:
        return serve_loop(jobs, budget=budget, slots=3, max_inflight=2, estimate=estimate, prepare=dev.prepare,
                          launch=dev.launch, submit=pool.submit, timeout=30)


def test_failures_stay_with_their_job():
    jobs = [_radial_job(i, arrival=0.001 * i) for i in range(12)]
    jobs[3].kspace = torch.zeros(4, 10, dtype=torch.complex64)  # malformed
    jobs[5].coords = np.zeros((2, 4, 16, 2), np.float32)       # stacked
    dev = _FakeDevice(fail_prepare={7}, fail_alloc={9})
    big = 11
    r = _run(jobs, dev, budget=10**6, estimate=lambda j: 10**7 if j.job_id == big else 1000)
    status = {j.job_id: j.status for j in jobs}
    assert {i for i, s in status.items() if s == "rejected"} == {3, 5, big}
    assert {i for i, s in status.items() if s == "failed"} == {7, 9}
    assert {i for i, s in status.items() if s == "done"} == {0, 1, 2, 4, 6, 8, 10}
    assert "injected" in jobs[7].error and "MemoryError" in jobs[9].error and "budget" in jobs[big].error
    assert all(j.t_done is not None for j in r["done"]) and all(j.t_done is None for j in r["failed"])
    launched = [i for b in dev.launched for i in b]
    assert sorted(launched) == [0, 1, 2, 4, 6, 8, 10]
    s = summarize(r, "fp32", 10**6)
    assert s["jobs"] == 7 and len(s["rejected"]) == 3 and len(s["failed"]) == 2 and s["p99_ms"] >= s["p50_ms"] > 0


def test_budget_released_by_failed_jobs():
    """A budget that holds one job at a time: failed jobs must give theirs
    back, or the loop would stall before the last jobs."""
    jobs = [_radial_job(i) for i in range(6)]
    dev = _FakeDevice(fail_prepare={1}, fail_alloc={2, 4})
    r = _run(jobs, dev, budget=1000, estimate=lambda j: 1000)
    assert [j.job_id for j in r["done"]] == [0, 3, 5]
    assert r["peak_used"] == 1000


def test_every_job_rejected_or_failed_ends_the_run():
    jobs = [_radial_job(i) for i in range(3)]
    jobs[0].kspace = torch.zeros(4, 1, dtype=torch.complex64)
    dev = _FakeDevice(fail_prepare={1}, fail_alloc={2})
    r = _run(jobs, dev)
    assert not r["done"] and len(r["rejected"]) == 1 and len(r["failed"]) == 2
    s = summarize(r, "bf16x3", 10**9)
    assert s["jobs"] == 0 and s["p99_ms"] == 0.0


def test_loop_timeout_raises_and_waits_for_preparations():
    class Never(_Handle):
        def query(self):
            return False

    dev = _FakeDevice()
    with cf.ThreadPoolExecutor(max_workers=2) as pool:
        with pytest.raises(RuntimeError, match="exceeded"):
            serve_loop([_radial_job(0)], budget=10**9, slots=2, max_inflight=1, estimate=lambda j: 1,
                       prepare=dev.prepare, launch=lambda b: (Never(0), None, {}), submit=pool.submit,
                       timeout=0.05)
