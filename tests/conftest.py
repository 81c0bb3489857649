import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "cfd-craft_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libcfdgpu.so")


@pytest.fixture(scope="session")
def gpu_available():
    import ctypes

    try:
        import cfdgpu

        cfdgpu.lib()
    except Exception as e:  # the library must exist on a GPU box; fail loudly there
        pytest.fail(f"libcfdgpu.so not loadable: {e}")
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_int(0)
    if hip.hipGetDeviceCount(ctypes.byref(n)) != 0 or n.value < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return True
