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
    """1s^2 plus, in every channel (l1, l2, L), a seeded complex multiple of the product of
    the lowest l1 state of charge 2 and the lowest l2 state of charge 1 (unnormalised): a
    smooth state that populates all channels."""
    from pyprop_amd.systems import helium_initial_state, hydrogenic_1s
    rng = np.random.default_rng(seed)
    basis = sysm.bases[1]
    psi = helium_initial_state(sysm)
    lmax = max(max(ch.l1, ch.l2) for ch in sysm.channels)
    inner = [hydrogenic_1s(basis, 2.0, l)[0] for l in range(lmax + 1)]
    outer = [hydrogenic_1s(basis, 1.0, l)[0] for l in range(lmax + 1)]
    for c, ch in enumerate(sysm.channels):
        a = amplitude * complex(rng.standard_normal(), rng.standard_normal())
        psi[c] += a * np.multiply.outer(inner[ch.l1], outer[ch.l2])
    return psi
