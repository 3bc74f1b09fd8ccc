import os
import sys

import pytest

# torch first: it loads the HIP runtime it was built with, which the library
# then shares; loaded after the library, torch finds no GPU
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """A cav_ctx on device 0. GPU tests fail (not skip) without a GPU."""
    from cavalry_amd import Context
    c = Context(0)
    yield c
    c.close()
