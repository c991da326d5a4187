"""Shared test set-up.  GPU tests carry @pytest.mark.gpu; everything else
runs on the CPU and never calls into the GPU library."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and runs the HIP kernels")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def fvc():
    """The GPU library; fails (never skips) when it cannot be loaded."""
    import fortranfvcell_amd as F
    F.init(0)
    return F


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
