import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libmplb.so")


def _hip_devices():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            continue
        n = ctypes.c_int(0)
        if lib.hipGetDeviceCount(ctypes.byref(n)) == 0:
            return n.value
        return 0
    return 0


def pytest_collection_modifyitems(config, items):
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    if _hip_devices() > 0:
        return
    skip = pytest.mark.skip(reason="no HIP device visible")
    for it in gpu_items:
        it.add_marker(skip)
