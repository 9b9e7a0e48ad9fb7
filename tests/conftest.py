import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle():
    """The CPU oracle (checker only), built on first use."""
    from oracle import oracle as O
    O.build()
    assert O.rounding_selftest(), "oracle directed rounding is broken on this host"
    return O


@pytest.fixture(scope="session")
def gpu():
    """libgpugo.so on device 0; the test fails (not skips) if it cannot load."""
    import gpugo_amd as G
    G.init(0)
    return G
