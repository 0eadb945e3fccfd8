import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def hip():
    """The HIP library; GPU tests fail loudly if it is missing."""
    from hidem_amd import _abi
    return _abi


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle
