import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the hip engine's kernels)")
    config.addinivalue_line("markers", "slow: full-size configuration (seconds of oracle time)")


@pytest.fixture(scope="session")
def lib():
    from vfhe_amd import _native
    return _native.load()


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    return O
