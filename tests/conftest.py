import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running check")


@pytest.fixture(scope="session")
def pkg():
    from util import load_pkg
    return load_pkg()


@pytest.fixture(scope="session")
def engine(pkg):
    eng = pkg.engine
    if eng.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return eng
