import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def _gpu_available():
    """True when the HIP runtime reports a device (asks the library itself)."""
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        cnt = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(cnt)) == 0 and cnt.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def yq():
    """The HIP library binding; builds it in-tree if it is missing."""
    import yannq_amd
    if not os.path.exists(yannq_amd.LIB_PATH):
        from yannq_amd import build
        build.build()
    yannq_amd.lib()
    return yannq_amd


@pytest.fixture()
def ctx(yq):
    if not _gpu_available():
        pytest.skip("no GPU visible")
    c = yq.Context(0)
    yield c
    c.close()
