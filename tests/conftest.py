import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available():
    """Whether this host exposes an AMD GPU (the kernel driver's device node).  No GPU runtime is
    loaded to find out."""
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def oracle():
    import oracle_marslib

    return oracle_marslib
