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
