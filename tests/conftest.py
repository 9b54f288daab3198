import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the libhns kernels")


@pytest.fixture(scope="session")
def hip():
    """GPU tests fail loudly (never skip) when the device or the HIP library is missing."""
    import torch

    import hnanosolver_amd as hns

    hns.lib()  # raises ImportError when libhns.so is not built
    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    return torch
