import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def gpu():
    """The product library with a usable device; fails (not skips) otherwise."""
    import ferray_amd
    assert ferray_amd.gpu_supports(), (
        "no usable gfx950 device: " + ferray_amd.lib().ferray_status_string(
            ferray_amd.lib().ferray_gemm_last_status()).decode())
    return ferray_amd
