"""The reference's Python binding (python/TensorBLAS.py:1-68), same functions and
meaning, over this library's C ABI: ``tc_syrk_wrapper`` and ``tc_trsm_wrapper``
loaded from ``../build/TensorBLAS.so`` next to this file.

Beyond the reference's two functions, every routine is exposed on torch tensors, in
every tier:

* ``tc_syr2k``, ``tc_trmm``, ``tc_symm`` follow ``tc_syrk`` / ``tc_trsm``: the
  reference routines' fixed option sets, returning a new tensor.  ``tc_syrk`` and
  ``tc_trsm`` take a ``tier`` too; the default "plain" keeps the reference's call.
* ``syrk``, ``syr2k``, ``trmm``, ``trsm``, ``symm`` take the BLAS options (side, uplo,
  trans, diag, alpha, beta); ``potrf`` factors and ``trtri`` inverts in place.

Tensors are 2-D CUDA tensors, float32 for the "plain", "split" and "fp32" tiers and
float64 for "fp64" and "ozaki".  A tensor may be row-major (torch's default) or a
column-major view (``t.T`` of a contiguous tensor); the call runs in the storage order
of its output, and an input stored the other way is copied into that order first.
Calls are queued on torch's current stream.
"""
import ctypes
import os

import torch

import tensorblas as _tb

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


def tc_syrk(A, transpose=False, nb=256, tier="plain"):
    """C = A @ A.T (full symmetric n x n) for an n x k tensor A; with transpose=True,
    C = A.T @ A."""
    if tier != "plain":
        C = syrk(A, trans="T" if transpose else "N", uplo="U", tier=tier, nb=nb)
        return _symmetrize(C, tier)
    if transpose:
        A = torch.clone(A.T).contiguous()
    A = A.contiguous()
    n, k = A.size(0), A.size(1)
    C = torch.zeros(n, n, device=A.device, dtype=torch.float32)
    status = lib.tc_syrk_wrapper(ctypes.c_long(n), ctypes.c_long(k), _fp(A), _fp(C), ctypes.c_long(nb))
    assert status == 0
    return C


def tc_trsm(A, B, nb=256, tier="plain"):
    """Solves X A^T = B for X (A n x n lower triangular, B m x n), returns X."""
    if tier != "plain":
        return trsm(A, B.clone(), side="R", uplo="L", trans="T", tier=tier, nb=nb)
    n = A.size(0)
    m = B.size(0)
    A = torch.clone(A.T).contiguous()
    B_inout = torch.clone(B.T).contiguous()
    status = lib.tc_trsm_wrapper(ctypes.c_long(m), ctypes.c_long(n), _fp(A), _fp(B_inout),
                                 ctypes.c_long(nb))
    assert status == 0
    return B_inout.T


def tc_syr2k(A, B, transpose=False, nb=256, tier="plain"):
    """C = A @ B.T + B @ A.T (full symmetric n x n) for n x k tensors A and B; with
    transpose=True, C = A.T @ B + B.T @ A."""
    C = syr2k(A, B, trans="T" if transpose else "N", uplo="U", tier=tier, nb=nb)
    return _symmetrize(C, tier)


def tc_trmm(A, B, nb=256, tier="plain"):
    """C = A @ B for A m x m lower triangular (its strict upper triangle is not read)
    and B m x n."""
    return trmm(A, B, side="L", uplo="L", trans="N", tier=tier, nb=nb)


def tc_symm(A, B, tier="plain"):
    """C = A @ B for A m x m symmetric with its lower triangle stored and B m x n."""
    return symm(A, B, side="L", uplo="L", tier=tier)


# ---------------------------------------------------------------- tensor calls
_handles = {}


def _handle(t):
    """the device's library handle, on torch's current stream"""
    dev = t.device.index if t.device.index is not None else torch.cuda.current_device()
    h = _handles.get(dev)
    if h is None:
        with torch.cuda.device(dev):
            h = _handles[dev] = _tb.Handle()
    h.set_stream(torch.cuda.current_stream(dev))
    return h


def _dtype(tier):
    if tier not in _tb.TIERS:
        raise ValueError(f"unknown tier {tier!r}")
    return torch.float64 if tier in ("fp64", "ozaki") else torch.float32


def _layout(t):
    """('R', ld) for row-major storage, ('C', ld) for column-major, None otherwise"""
    r, c = t.shape
    s0, s1 = t.stride()
    if s1 == 1 and s0 >= max(1, c):
        return "R", s0
    if s0 == 1 and s1 >= max(1, r):
        return "C", s1
    return None


def _check(tier, *ts):
    dt = _dtype(tier)
    for t in ts:
        if not isinstance(t, torch.Tensor) or t.dim() != 2:
            raise ValueError("operands are 2-D torch tensors")
        if not t.is_cuda:
            raise ValueError("operands are CUDA tensors")
        if t.dtype != dt:
            raise ValueError(f"tier {tier!r} takes {dt} tensors, got {t.dtype}")
        if t.device != ts[0].device:
            raise ValueError("operands are on one device")


def _out_layout(t):
    lay = _layout(t)
    if lay is None:
        raise ValueError("the output tensor is neither row- nor column-major")
    return lay


def _as(t, order):
    """(t in the given storage order, its leading dimension); copies when needed"""
    lay = _layout(t)
    if lay is None or lay[0] != order:
        t = t.contiguous() if order == "R" else t.T.contiguous().T
        lay = _layout(t)
    return t, lay[1]


def _new(rows, cols, like):
    return torch.zeros(rows, cols, device=like.device, dtype=like.dtype)


def _symmetrize(C, tier):
    """copies the upper triangle of the contiguous (row-major) C onto its lower one: in
    column-major terms, the lower triangle of the same memory onto its upper"""
    _handle(C).copy_lower_to_upper(int(_dtype(tier) == torch.float64), C.shape[0], C.data_ptr(),
                                   C.stride(0))
    return C


def _dims_square(A, n, name):
    if A.shape != (n, n):
        raise ValueError(f"{name} is {n} x {n}, got {tuple(A.shape)}")


def trsm(A, B, side="L", uplo="L", trans="N", diag="N", alpha=1.0, tier="split", nb=256):
    """op(A) X = alpha B (side L) or X op(A) = alpha B (side R), A triangular; B is
    overwritten with X and returned."""
    _check(tier, A, B)
    m, n = B.shape
    _dims_square(A, m if side.upper() == "L" else n, "A")
    order, ldb = _out_layout(B)
    A, lda = _as(A, order)
    _handle(B).trsm_ex(tier, side, uplo, trans, diag, m, n, alpha, A.data_ptr(), lda, B.data_ptr(),
                       ldb, nb=nb, order=order)
    return B


def trmm(A, B, side="L", uplo="L", trans="N", diag="N", alpha=1.0, tier="split", nb=256, out=None):
    """alpha op(A) B (side L) or alpha B op(A) (side R), A triangular, into out (a new
    tensor when None)."""
    _check(tier, A, B, *(() if out is None else (out,)))
    m, n = B.shape
    _dims_square(A, m if side.upper() == "L" else n, "A")
    C = _new(m, n, B) if out is None else out
    if C.shape != (m, n):
        raise ValueError(f"out is {m} x {n}, got {tuple(C.shape)}")
    order, ldc = _out_layout(C)
    A, lda = _as(A, order)
    B, ldb = _as(B, order)
    _handle(C).trmm_ex(tier, side, uplo, trans, diag, m, n, alpha, A.data_ptr(), lda, B.data_ptr(),
                       ldb, C.data_ptr(), ldc, nb=nb, order=order)
    return C


def symm(A, B, C=None, side="L", uplo="L", alpha=1.0, beta=0.0, tier="split"):
    """C = alpha A B + beta C (side L) or alpha B A + beta C (side R), A symmetric with
    its uplo triangle read; C is a new tensor when None."""
    _check(tier, A, B, *(() if C is None else (C,)))
    m, n = B.shape
    _dims_square(A, m if side.upper() == "L" else n, "A")
    if C is None:
        C, beta = _new(m, n, B), 0.0
    if C.shape != (m, n):
        raise ValueError(f"C is {m} x {n}, got {tuple(C.shape)}")
    order, ldc = _out_layout(C)
    A, lda = _as(A, order)
    B, ldb = _as(B, order)
    _handle(C).symm_ex(tier, side, uplo, m, n, alpha, A.data_ptr(), lda, B.data_ptr(), ldb, beta,
                       C.data_ptr(), ldc, order=order)
    return C


def syrk(A, C=None, uplo="L", trans="N", alpha=1.0, beta=0.0, tier="split", nb=256):
    """C = alpha op(A) op(A)^T + beta C, op(A) = A (n x k) or A^T; only the uplo
    triangle of C is written (C is a new zero tensor when None)."""
    _check(tier, A, *(() if C is None else (C,)))
    n, k = A.shape if trans.upper() == "N" else A.shape[::-1]
    if C is None:
        C, beta = _new(n, n, A), 0.0
    _dims_square(C, n, "C")
    order, ldc = _out_layout(C)
    A, lda = _as(A, order)
    _handle(C).syrk_ex(tier, uplo, trans, n, k, alpha, A.data_ptr(), lda, beta, C.data_ptr(), ldc,
                       nb=nb, order=order)
    return C


def syr2k(A, B, C=None, uplo="L", trans="N", alpha=1.0, beta=0.0, tier="split", nb=256):
    """C = alpha (op(A) op(B)^T + op(B) op(A)^T) + beta C; only the uplo triangle of C
    is written (C is a new zero tensor when None)."""
    _check(tier, A, B, *(() if C is None else (C,)))
    if A.shape != B.shape:
        raise ValueError("A and B have one shape")
    n, k = A.shape if trans.upper() == "N" else A.shape[::-1]
    if C is None:
        C, beta = _new(n, n, A), 0.0
    _dims_square(C, n, "C")
    order, ldc = _out_layout(C)
    A, lda = _as(A, order)
    B, ldb = _as(B, order)
    _handle(C).syr2k_ex(tier, uplo, trans, n, k, alpha, A.data_ptr(), lda, B.data_ptr(), ldb, beta,
                        C.data_ptr(), ldc, nb=nb, order=order)
    return C


def potrf(A, upper=False, tier="split", nb=256):
    """Cholesky factorization in place: A = L L^T with L in the lower triangle
    (upper=False) or A = U^T U with U in the upper triangle; the other triangle is
    neither read nor written.  Returns LAPACK's info (0, or the order of the first
    leading minor that is not positive definite); synchronizes the stream."""
    _check(tier, A)
    n = A.shape[0]
    _dims_square(A, n, "A")
    order, lda = _out_layout(A)
    # the library factors the column-major lower triangle; row-major upper is the same
    # memory (U = L^T)
    if (order == "C") != (not upper):
        # the other triangle: factor a copy stored the other way, then copy back
        At = A.T.contiguous().T if order == "R" else A.contiguous()
        info = potrf(At, upper=upper, tier=tier, nb=nb)
        A.copy_(At)
        return info
    return _handle(A).potrf(tier, n, A.data_ptr(), lda, nb=nb)


def trtri(A, upper=False, unit=False, tier="split", nb=256):
    """Triangular inverse in place: the lower (upper=False) or upper triangle of A is
    replaced by that of inv(A); the other strict triangle (and, with unit=True, the
    diagonal) is neither read nor written.  Returns LAPACK's info (0, or i for a zero
    A(i, i), A then unchanged); synchronizes the stream."""
    _check(tier, A)
    n = A.shape[0]
    _dims_square(A, n, "A")
    order, lda = _out_layout(A)
    # a row-major triangle is the other triangle of the column-major memory, and
    # inv(A^T) = inv(A)^T, so the call is the column-major one with uplo exchanged
    uplo = "U" if upper == (order == "C") else "L"
    return _handle(A).trtri(tier, uplo, "U" if unit else "N", n, A.data_ptr(), lda, nb=nb)
