import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU test")


def _ensure_built():
    """Build the oracle and the HIP library in-tree if they are missing."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if not os.path.exists(os.path.join(ROOT, "libatssim_amd", "lib", "libatssim_hip.so")):
        subprocess.check_call(["make", "-s", "-j8", "-C", os.path.join(ROOT, "libatssim_amd", "csrc")])


_ensure_built()


@pytest.fixture(scope="session")
def gpu_available():
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        ok = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        ok = False
    if not ok:
        pytest.skip("no HIP device")
    return True


def pytest_sessionfinish(session, exitstatus):
    """Destroy every GPU context the tests left open while the HIP runtime is
    still up (objects left to garbage collection at interpreter exit may be
    destroyed after it has shut down)."""
    mod = sys.modules.get("libatssim_amd._lib")
    if mod is not None:
        mod.close_all()
