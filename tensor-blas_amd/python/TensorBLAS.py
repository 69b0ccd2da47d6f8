"""The reference's Python binding (python/TensorBLAS.py:1-68), same functions and
meaning, over this library's C ABI: ``tc_syrk_wrapper`` and ``tc_trsm_wrapper``
loaded from ``../build/TensorBLAS.so`` next to this file.
"""
import ctypes
import os

import torch

dir_path = os.path.dirname(os.path.realpath(__file__))

lib = ctypes.CDLL(os.path.join(dir_path, "../build/TensorBLAS.so"))
lib.tc_syrk_wrapper.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float), ctypes.c_long]
lib.tc_syrk_wrapper.restype = ctypes.c_int
lib.tc_trsm_wrapper.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float), ctypes.c_long]
lib.tc_trsm_wrapper.restype = ctypes.c_int


def _fp(t):
    return ctypes.cast(t.data_ptr(), ctypes.POINTER(ctypes.c_float))


def tc_syrk(A, transpose=False, nb=256):
    """C = A @ A.T (full symmetric n x n) for a row-major n x k float32 tensor A;
    with transpose=True, C = A.T @ A."""
    if transpose:
        A = torch.clone(A.T).contiguous()
    A = A.contiguous()
    n, k = A.size(0), A.size(1)
    C = torch.zeros(n, n, device=A.device, dtype=torch.float32)
    status = lib.tc_syrk_wrapper(ctypes.c_long(n), ctypes.c_long(k), _fp(A), _fp(C), ctypes.c_long(nb))
    assert status == 0
    return C


def tc_trsm(A, B, nb=256):
    """Solves X A^T = B for X (A n x n lower triangular, B m x n), returns X."""
    n = A.size(0)
    m = B.size(0)
    A = torch.clone(A.T).contiguous()
    B_inout = torch.clone(B.T).contiguous()
    status = lib.tc_trsm_wrapper(ctypes.c_long(m), ctypes.c_long(n), _fp(A), _fp(B_inout),
                                 ctypes.c_long(nb))
    assert status == 0
    return B_inout.T
