import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (HERE, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle_bind import Oracle
    return Oracle()


def gpu_available() -> bool:
    try:
        from hercules_amd import hkg
        n = hkg.lib().hkg_device_sync()
        return n == 0
    except Exception:
        return False
