import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def gpu_present():
    from neuro_edit_v2_amd import _lib
    return bool(_lib.lib().nh_available())


@pytest.fixture(scope="session")
def require_gpu():
    if not gpu_present():
        pytest.fail("gpu test without a usable HIP device")
