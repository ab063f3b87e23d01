import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libcasim_dyad.so")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build(ref=os.path.isdir("/root/reference/CaSim"))
    return oracle


@pytest.fixture(scope="session")
def ref_lib(oracle_mod):
    lib = oracle_mod.load_ref()
    if lib is None:
        pytest.skip("oracle/_ref (the reference's own RNG/kinetic sources) is not built here")
    return lib


@pytest.fixture(scope="session")
def gpu():
    from casim_amd import abi
    lib = abi.load()
    import ctypes
    n = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount(ctypes.byref(n))
    except OSError:
        pass
    if n.value < 1:
        pytest.fail("GPU test run without a HIP device")
    return lib
