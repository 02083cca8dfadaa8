import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture
def rng():
    return np.random.default_rng(1234)


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)



def helium_smooth_start(sysm, seed=17, amplitude=0.05):
    from pyprop_amd.systems import helium_smooth_start as f
    return f(sysm, seed, amplitude)
