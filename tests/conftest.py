import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through libhzo.so")


def _gpu_present() -> bool:
    import cmsl_amd
    return cmsl_amd.device_count() > 0


def pytest_collection_modifyitems(config, items):
    gpu_items = [it for it in items if "gpu" in it.keywords]
    if not gpu_items:
        return
    # libhzo.so must load on every machine; a missing build is an error, never a skip.
    import cmsl_amd
    cmsl_amd.lib()
    if _gpu_present():
        return
    if os.path.exists("/dev/kfd"):
        # A GPU node whose device is not visible must fail loudly, not skip.
        return
    skip = pytest.mark.skip(reason="no HIP device in this container")
    for it in gpu_items:
        it.add_marker(skip)
