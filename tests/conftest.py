import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: longer CPU checks")


@pytest.fixture(scope="session")
def cu():
    """The package (directory chaos-ultra_amd)."""
    return importlib.import_module("chaos-ultra_amd")


@pytest.fixture(scope="session")
def oracle():
    import _oracle
    _oracle.lib()
    return _oracle


@pytest.fixture(scope="session")
def gpu(cu):
    """Initialise device 0; fails (does not skip) when no device is visible."""
    cu.init(0)
    return cu


@pytest.fixture(scope="session")
def engine(gpu):
    e = gpu.Engine()
    yield e
    e.close()


GOLDEN = os.path.join(TESTS, "golden")
