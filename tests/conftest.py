import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); the parity tests proper")


@pytest.fixture(autouse=True, scope="session")
def _oracle_rootfinder():
    """The oracle uses the reference's own rpoly (oracle/_ref) when it is built, else its Aberth solver."""
    import oracle

    which = oracle.use_default_rootfinder()
    print(f"oracle root finder: {which}")
    yield


@pytest.fixture(scope="session")
def core_harness():
    """Host build of the product core (test-only), compiled on demand."""
    import ctypes
    import subprocess

    so = os.path.join(ROOT, "tests", "_build", "core_harness.so")
    src = os.path.join(ROOT, "tests", "native", "core_harness.cpp")
    hdrs = [os.path.join(ROOT, "collisiondetection_amd", "csrc", "ccd_core.h")]
    newest = max(os.path.getmtime(x) for x in [src] + hdrs)
    if not (os.path.exists(so) and os.path.getmtime(so) >= newest):
        subprocess.run(["bash", os.path.join(ROOT, "tests", "native", "build.sh")], check=True)
    return ctypes.CDLL(so)
