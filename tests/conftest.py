import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and libpsfun_hip.so")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def gpu_ctx():
    """The default libpsfun_hip context on cuda:0; GPU tests fail loudly
    (never fall back) when the library or the device is missing."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a visible HIP device")
    from psfun_amd import psb
    return psb.default_context(0)
