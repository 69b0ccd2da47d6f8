"""The reference's C++ entry points called from C++ (include/TensorBLAS.h:139-167 of
the reference, plus tc_ozimmu_symm): tests/cpp/tc_boundary.cpp calls each of the 15
tc_* functions as the reference's testing/ drivers do and compares its result bit for
bit with the tb_* C-ABI call of the same routine and tier.  Sizes, leading dimensions
and alpha/beta are all distinct, so a forwarder that passes an argument in the wrong
place fails."""
import ctypes
import os

import pytest

from _tb import ROOT

BOUNDARY_LIB = os.path.join(ROOT, "tests", "cpp", "build", "libtc_boundary.so")
REFERENCE_CPP = ["tc_syrk", "tc_cumpsgemm_syrk", "tc_ozimmu_syrk", "tc_trsm", "tc_cumpsgemm_trsm",
                 "tc_ozimmu_trsm", "tc_trmm", "tc_cumpsgemm_trmm", "tc_ozimmu_trmm", "tc_syr2k",
                 "tc_cumpsgemm_syr2k", "tc_ozimmu_syr2k", "tc_symm", "tc_cumpsgemm_symm",
                 "tc_ozimmu_symm"]


def _load():
    if not os.path.exists(BOUNDARY_LIB):
        pytest.skip("tests/cpp/build/libtc_boundary.so not built (__graft_entry__.build())")
    L = ctypes.CDLL(BOUNDARY_LIB)
    L.tc_boundary_check.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.tc_boundary_check.restype = ctypes.c_int
    return L


def test_boundary_caller_loads_and_links_the_library():
    """CPU: the caller library loads (resolving every tc_* it calls from libTensorBLAS.so)
    and exports its entry point; nothing is computed."""
    L = _load()
    assert hasattr(L, "tc_boundary_check")


@pytest.mark.gpu
def test_every_tc_entry_point_equals_its_tb_call():
    L = _load()
    buf = ctypes.create_string_buffer(8192)
    checked = ctypes.c_int(0)
    failures = L.tc_boundary_check(buf, len(buf), ctypes.byref(checked))
    report = buf.value.decode()
    print(report)
    assert failures == 0, report
    assert checked.value == 15, report
    for name in REFERENCE_CPP:
        assert f"{name}: tc status 0, tb status 0, bitwise equal" in report, (name, report)
