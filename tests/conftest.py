import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def gpu_available():
    try:
        import torch  # noqa: F401  (only to count devices; no allocation)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_ctx():
    from usbs_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libusbs_hip.so is not built: run python -m usbs_amd.build")
    return _lib.Context(0)
