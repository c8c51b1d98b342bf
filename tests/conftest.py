import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly when a GPU test runs without a GPU: there is no CPU fallback."""
    from cfs_plume_amd import hip
    if not hip.gpu_available():
        pytest.fail("no HIP device visible: GPU tests need an MI355X")
    return True
