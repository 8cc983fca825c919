import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libplato_hip.so)")


@pytest.fixture(scope="session")
def ctx():
    """One device context for the GPU tests; fails (not skips) when the library is absent."""
    from plato_amd import Context
    c = Context(0)
    yield c
    c.close()


def golden(name):
    return os.path.join(GOLDEN, name)
