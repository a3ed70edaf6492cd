import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")
    config.addinivalue_line("markers", "slow: large inputs (tens of seconds)")


@pytest.fixture(scope="session")
def gk():
    from _pkg import gk as mod
    return mod


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    return _oracle


@pytest.fixture(scope="session")
def ctx(gk):
    c = gk.Context(0)
    yield c
    c.close()
