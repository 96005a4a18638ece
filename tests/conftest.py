import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"
REF_DATA = GOLDEN / "ref_data"
SYNTH = GOLDEN / "synth"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); runs the HIP search path")


@pytest.fixture(scope="session")
def oracle():
    from oracle.refpy import Oracle

    o = Oracle()
    o.set_threads(8)
    return o


@pytest.fixture(scope="session")
def reference():
    from oracle.refpy import REF_SO, Reference

    if not REF_SO.exists():
        pytest.skip("the reference library (oracle/_ref) is not built here")
    return Reference()


def has_gpu():
    try:
        import ctypes

        lib = ctypes.CDLL(str(ROOT / "qann_amd" / "libqann_hip.so"))
        n = ctypes.c_int(0)
        return lib.qann_hip_device_count(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False
