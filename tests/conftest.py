"""Shared fixtures.  GPU tests carry ``@pytest.mark.gpu`` and run on a HIP
device; everything else runs on the CPU (oracle, host logic, ABI)."""

import os
import pathlib
import sys

import numpy as np
import pytest
import torch

ROOT = pathlib.Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP (MI355X) device and the built HIP library")
    config.addinivalue_line("markers", "slow: larger problem sizes")


def has_gpu() -> bool:
    return torch.cuda.is_available()


@pytest.fixture
def rng():
    return np.random.default_rng(42)


@pytest.fixture
def torch_rng():
    return torch.Generator().manual_seed(42)


@pytest.fixture
def device():
    """The HIP device (GPU tests only)."""
    if not has_gpu():
        pytest.skip("no HIP device")
    return torch.device("cuda")


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case before any GPU step, so a fault report names its test
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[case] {request.node.nodeid}", flush=True)
    yield


_ = os
