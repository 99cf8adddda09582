import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity against the CPU oracle")


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library with a visible device; fails (does not skip) otherwise."""
    import akaze_amd
    n = akaze_amd.device_count()
    assert n >= 1, "no HIP device visible: GPU tests need an MI355X"
    return akaze_amd
