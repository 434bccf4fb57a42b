import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    config.addinivalue_line("markers", "slow: full-size workloads")
    config.addinivalue_line("markers", "fallback: may use the library's opt-in CPU path (error-path tests only)")


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


@pytest.fixture(autouse=True)
def device_results_only(request):
    """A GPU test's results must come from the device: unless the test is an error-path test (marked fallback), the
    library's CPU path stays off and its use counter does not move."""
    if request.node.get_closest_marker("gpu") is None or request.node.get_closest_marker("fallback") is not None:
        yield
        return
    import kolmo_amd
    kolmo_amd.set_cpu_fallback(False)
    kolmo_amd.fault_inject(None)
    before = kolmo_amd.counters()["cpu_fallbacks"]
    yield
    assert kolmo_amd.counters()["cpu_fallbacks"] == before, "a GPU test's result came from the CPU path"
