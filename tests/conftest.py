import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "check_n200.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def capi():
    from lf_dem_amd import capi as C
    C.lib()
    if C.lib().lfd_device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return C


_CONF_CACHE = {}


def conf(n, phi, seed):
    """Seeded dense bidisperse configuration (cached per session)."""
    key = (n, phi, seed)
    if key not in _CONF_CACHE:
        from lf_dem_amd import confgen
        _CONF_CACHE[key] = confgen.generate(n, phi, seed)
    return _CONF_CACHE[key]
