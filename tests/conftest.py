import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU test")


def gpu_available():
    try:
        from pysweep_amd import hip
        return hip.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def device():
    """Initialise device 0 through the C-ABI library; fail (not skip) when the
    library is missing, so a GPU run never passes without native code."""
    from pysweep_amd import hip
    hip.init(0)
    return hip
