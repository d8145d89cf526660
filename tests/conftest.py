import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mpi-multi-process-version_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, PKG)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (MI355X, gfx950)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def oracle_mod():
    import lbm_oracle
    lbm_oracle.lib()
    lbm_oracle.set_threads(min(8, os.cpu_count() or 1))
    return lbm_oracle


@pytest.fixture(scope="session")
def hip():
    import lbmhip
    lbmhip.load()
    return lbmhip
