import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu_session_factory():
    """Factory of GPU sessions; fails loudly if the native library is absent."""
    from ifem_amd import Session, lib
    lib()  # raises if libifem_gpu.so was not built

    def make():
        return Session()
    return make


@pytest.fixture(scope="session", autouse=True)
def torch_runtime_first():
    """torch ships its own HIP runtime next to the system one that
    libifem_gpu.so links; when both are loaded in one process, torch's must
    open the device first, or torch reports no GPU afterwards.  GPU tests that
    hand torch tensors to the library therefore start torch's runtime at the
    beginning of the session (counting devices does not initialise it)."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.device_count() > 0:
        torch.cuda.init()
