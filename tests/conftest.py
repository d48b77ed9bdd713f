import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "janus-sim_amd"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


def gpu_available() -> bool:
    try:
        import janus_treepm as j
        return j.device_count() > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle_lib
    return oracle_lib.load()


@pytest.fixture(scope="session")
def jtp():
    import janus_treepm as j
    if j.device_count() <= 0:
        pytest.fail("GPU test selected but no HIP device is visible")
    return j
