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
