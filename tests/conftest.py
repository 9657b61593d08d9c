"""Shared test setup: the package directory on sys.path, the `gpu` marker, seeded states."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
PACKAGE_DIR = ROOT / "rcmpscore.jl_amd"
for p in (str(ROOT), str(PACKAGE_DIR)):
    if p not in sys.path:
        sys.path.insert(0, p)

FIXTURES = Path(__file__).resolve().parent / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def random_state_arrays(d, seed, scaling=1.0):
    """Complex-normal R, K (K Hermitised) in the draw order of RandomState (src/state.jl:55-60):
    the matrix that becomes K first, then R.  numpy PCG64 stands in for Julia's MersenneTwister."""
    rng = np.random.default_rng(seed)
    def crandn():
        return (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2)
    k = crandn()
    r = scaling * crandn()
    return r, scaling * (k + k.conj().T)


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
