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
    so_libm = os.path.join(ROOT, "tests", "_build", "core_harness_libm.so")
    src = os.path.join(ROOT, "tests", "native", "core_harness.cpp")
    hdrs = [os.path.join(ROOT, "collisiondetection_amd", "csrc", h) for h in ("ccd_core.h", "ccd_rpoly.h")]
    newest = max(os.path.getmtime(x) for x in [src] + hdrs)
    if not all(os.path.exists(x) and os.path.getmtime(x) >= newest for x in (so, so_libm)):
        subprocess.run(["bash", os.path.join(ROOT, "tests", "native", "build.sh")], check=True)
    return ctypes.CDLL(so)


@pytest.fixture(scope="session")
def core_harness_libm(core_harness):
    """The same host build with the C library's log/exp in the reference-root solver (test-only)."""
    import ctypes

    return ctypes.CDLL(os.path.join(ROOT, "tests", "_build", "core_harness_libm.so"))
