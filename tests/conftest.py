import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: full-size workloads")


def gpu_ready():
    """True when the HIP library loads and a gfx950 device answers."""
    try:
        import kolmo_amd
        return kolmo_amd.device_ok()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not gpu_ready():
        pytest.fail("no usable gfx950 device or libkbwt_hip.so is missing: GPU tests cannot run")
    import kolmo_amd
    return kolmo_amd
