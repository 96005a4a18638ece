"""The C ABI library loads and exports every entry point include/qann_hip.h
declares (no GPU needed: nothing here launches work)."""
import ctypes
import re

from capi import IvfDesc
from conftest import ROOT


def declared():
    text = (ROOT / "include" / "qann_hip.h").read_text()
    return sorted(set(re.findall(r"\b(qann_hip_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("qann_hip_ivf_create", "qann_hip_ivf_set_lists", "qann_hip_search", "qann_hip_search_async",
                 "qann_hip_refine_create", "qann_hip_flat_create", "qann_hip_set_ids", "qann_hip_index_destroy"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "qann_amd" / "libqann_hip.so"))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_is_empty_string_initially():
    lib = ctypes.CDLL(str(ROOT / "qann_amd" / "libqann_hip.so"))
    lib.qann_hip_last_error.restype = ctypes.c_char_p
    assert lib.qann_hip_last_error() == b""


def test_invalid_arguments_are_rejected_before_any_device_work():
    lib = ctypes.CDLL(str(ROOT / "qann_amd" / "libqann_hip.so"))
    lib.qann_hip_last_error.restype = ctypes.c_char_p
    out = ctypes.c_void_p()
    # nlist outside [100, 65535] is the reference's std::invalid_argument
    d = IvfDesc(dim=16, nlist=50, metric=0, m=8)
    assert lib.qann_hip_ivf_create(ctypes.byref(d), ctypes.byref(out)) == 1
    assert b"nlist" in lib.qann_hip_last_error()
    d = IvfDesc(dim=16, nlist=100, metric=0, m=3)
    assert lib.qann_hip_ivf_create(ctypes.byref(d), ctypes.byref(out)) == 1
    # SDC with fast-scan lists (ivf_index.hpp:95-96)
    d = IvfDesc(dim=16, nlist=100, metric=0, m=8, storage=0, pq_sdc=1)
    assert lib.qann_hip_ivf_create(ctypes.byref(d), ctypes.byref(out)) == 1
    assert b"SDC" in lib.qann_hip_last_error()
