import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "classicalspinmc.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcsmc on the device)")


@pytest.fixture
def rng():
    return np.random.default_rng(12345)


def random_spins(N, S, rng):
    """uniform on the sphere of radius S, shape (N, 3)"""
    v = rng.normal(size=(N, 3))
    return S * v / np.linalg.norm(v, axis=1, keepdims=True)
