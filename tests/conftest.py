import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

# torch first: its bundled HIP runtime then serves libnca_hip as well (one runtime per
# process, whichever library is loaded first provides libamdhip64.so.7).
import torch  # noqa: E402,F401


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def gpu_available():
    try:
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    """Initialise libnca_hip on cuda:0 (GPU tests only)."""
    if not gpu_available():
        pytest.skip("no GPU")
    from ncarray_amd import capi
    capi.init(0)
    return torch.device("cuda:0")
