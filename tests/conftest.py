import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: longer CPU test")


def gpu_available():
    try:
        from channel_flow_amd import runtime
        return runtime.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def require_gpu():
    from channel_flow_amd import runtime
    n = runtime.device_count()
    if n <= 0:
        pytest.fail("gpu test selected but no HIP device is visible")
    return n
