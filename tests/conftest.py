import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)


def load_pkg():
    if "gwfa_amd" in sys.modules:
        return sys.modules["gwfa_amd"]
    pdir = os.path.join(ROOT, "gwfa-cuda-version_amd")
    spec = importlib.util.spec_from_file_location("gwfa_amd", os.path.join(pdir, "__init__.py"),
                                                  submodule_search_locations=[pdir])
    m = importlib.util.module_from_spec(spec)
    sys.modules["gwfa_amd"] = m
    spec.loader.exec_module(m)
    return m


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests through the C-ABI")


def gpu_present():
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def gw():
    return load_pkg()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(gw):
    """The GPU path; on a GPU box it must be built and loadable (no fallback)."""
    if not gpu_present():
        pytest.fail("gpu test selected but no HIP device is visible")
    gw.hip_lib()
    return gw


def read_ints(path):
    with open(path) as f:
        return [int(x) for x in f.read().split()]
