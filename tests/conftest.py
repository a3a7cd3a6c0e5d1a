import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# a batched search that has not finished by then stops with an error
os.environ.setdefault("PFVS_SEARCH_DEADLINE_S", "60")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def gpu_present():
    try:
        from pfvs_amd import gpu
        return gpu.available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def need_gpu():
    """GPU tests fail loudly (not skip) when the device or the HIP library is
    missing: a GPU run must exercise the native path."""
    from pfvs_amd import gpu
    if not gpu.available():
        pytest.fail("HIP backend unavailable: " + gpu.unavailable_reason())
    return gpu
