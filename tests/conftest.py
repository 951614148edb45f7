import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(autouse=True)
def oracle_threads(request):
    """The oracle's OpenMP thread count is process-wide: GPU tests compare
    against large oracle runs and take up to 16 threads; CPU tests run small
    grids, where threads cost more than they give (and several test workers
    may share the CPUs), so they take one."""
    from oracle import oracle as orc
    n = min(16, os.cpu_count() or 1) if request.node.get_closest_marker("gpu") else 1
    orc.set_threads(n)
    yield
    orc.set_threads(1)
