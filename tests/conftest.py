import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        from psp_amd import _lib
        return _lib.lib().psp_hip_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rand_block(rng, n, k):
    return np.ascontiguousarray(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))
