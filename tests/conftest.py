import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

# torch ships its own libamdhip64 (soname libamdhip64.so.7).  Loaded first, it
# also serves libgkway.so's dependency on that soname, so the process holds
# one HIP runtime; loaded after libgkway.so it would bring a second runtime
# into the process, under which torch may not see the device.
try:
    import torch  # noqa: F401
except ImportError:
    pass


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
