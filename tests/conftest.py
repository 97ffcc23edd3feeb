import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdogpup_hip.so)")


def small_slab(seed=1, jitter=0.2, extent=(20.0, 20.0, 10.0), h=2.0, mua=0.01, inclusions=()):
    from dogpup_amd import meshgen
    return meshgen.slab(extent, h, mua, 1.0, 1.4, jitter=jitter, seed=seed, inclusions=inclusions)


@pytest.fixture(scope="session")
def tiny_mesh():
    return small_slab()


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
