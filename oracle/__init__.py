"""Test-only CPU oracle for the Tensor-BLAS hot path.

This package is test infrastructure: only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import it, and only as the checker.  The
product library never loads it.

It wraps ``oracle/oracle.c`` (built to ``oracle/build/liboracle.so`` by
``make -C oracle``) with numpy helpers.  Every routine works on Fortran-ordered
float64 arrays (column-major, as the reference's device buffers).  Each C function
cites the reference file:line it restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_p = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: run `make -C {_HERE}` ({_LIB_PATH} missing)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_find_mat_size_syrk.argtypes = [_i64, _p]
        L.orc_find_mat_size_trsm.argtypes = [_i64, _p]
        L.orc_f32_to_f16.argtypes = [ctypes.c_float]
        L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_f16_to_f32.restype = ctypes.c_float
        L.orc_f32_to_bf16.argtypes = [ctypes.c_float]
        L.orc_f32_to_bf16.restype = ctypes.c_uint16
        L.orc_f32_to_f16_array.argtypes = [_i64, _p, _p]
        L.orc_round_f16_array.argtypes = [_i64, _p, _p]
        L.orc_dgemm.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _i64, _i64, _dbl, _p, _i64,
                                _p, _i64, _dbl, _p, _i64]
        L.orc_dsyrk_ln.argtypes = [_i64, _i64, _dbl, _p, _i64, _dbl, _p, _i64]
        L.orc_dsyr2k_ln.argtypes = [_i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64]
        L.orc_dtrmm_llnn.argtypes = [_i64, _i64, _dbl, _p, _i64, _p, _i64, _p, _i64]
        L.orc_dtrsm_rltn.argtypes = [_i64, _i64, _p, _i64, _p, _i64]
        L.orc_dtrsm_llnn.argtypes = [_i64, _i64, _p, _i64, _p, _i64]
        L.orc_dsymm_ll.argtypes = [_i64, _i64, _dbl, _p, _i64, _p, _i64, _dbl, _p, _i64]
        L.orc_tc_trsm_plain_emul.argtypes = [_i64, _i64, _p, _i64, _p, _i64, _i64, _p]
        L.orc_ozaki_slice.argtypes = [ctypes.c_int, _i64, _i64, _p, _i64, ctypes.c_int, _p, _p]
        L.orc_ozaki_gemm.argtypes = [_i64, _i64, _i64, ctypes.c_int, _p, _p, _p, _p, _dbl,
                                     _dbl, _p, _i64, _i64, ctypes.c_int]
        L.orc_errors.argtypes = [_i64, _i64, _p, _i64, _p, _i64, _p, _i64, ctypes.c_int, _p]
        L.orc_dpotrf_lower.argtypes = [_i64, _p, _i64]
        L.orc_rowdot_xp.argtypes = [_i64, _i64, _p, _i64, _p, _i64, _p, _i64, _p, _i64, _i64,
                                    _p, _p]
        L.orc_rowdot_xp.restype = None
        L.orc_dpotrf_lower.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


# ------------------------------------------------------------------ planners
def find_mat_size_syrk(n):
    """util/mat_size.cu:5-29 -> list of chunks, largest first (reference order)."""
    out = np.zeros(16, dtype=np.int64)
    c = lib().orc_find_mat_size_syrk(int(n), _ptr(out))
    if c < 0:
        raise ValueError(f"n={n} is outside the reference planner's domain")
    return [int(v) for v in out[:c]]


def find_mat_size_trsm(n):
    """util/mat_size.cu:31-55."""
    out = np.zeros(16, dtype=np.int64)
    c = lib().orc_find_mat_size_trsm(int(n), _ptr(out))
    if c < 0:
        raise ValueError(f"n={n} is outside the reference planner's domain")
    return [int(v) for v in out[:c]]


# --------------------------------------------------------------- conversions
def f32_to_f16_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().orc_f32_to_f16_array(x.size, _ptr(x), _ptr(out))
    return out


def round_f16(x):
    """FP32 -> FP16 (RNE) -> FP64, elementwise, keeping the memory order of x."""
    xf = np.asarray(x, dtype=np.float32)
    order = "F" if xf.flags.f_contiguous and not xf.flags.c_contiguous else "C"
    xc = np.require(xf, requirements=[order[0]])
    out = np.empty(xc.shape, dtype=np.float64, order=order)
    lib().orc_round_f16_array(xc.size, _ptr(xc), _ptr(out))
    return out


# ------------------------------------------------------------------- BLAS-3
def dgemm(A, B, transa=False, transb=False, alpha=1.0, beta=0.0, C=None):
    A = _f64(A); B = _f64(B)
    m = A.shape[1] if transa else A.shape[0]
    k = A.shape[0] if transa else A.shape[1]
    n = B.shape[0] if transb else B.shape[1]
    C = np.zeros((m, n), order="F") if C is None else _f64(C).copy(order="F")
    lib().orc_dgemm(int(transa), int(transb), m, n, k, alpha, _ptr(A), max(1, A.shape[0]),
                    _ptr(B), max(1, B.shape[0]), beta, _ptr(C), max(1, m))
    return C


def dsyrk(A, alpha=1.0, beta=0.0, C=None):
    """Lower triangle of alpha*A*A^T + beta*C (upper part of the returned C is C's input)."""
    A = _f64(A)
    n, k = A.shape
    C = np.zeros((n, n), order="F") if C is None else _f64(C).copy(order="F")
    lib().orc_dsyrk_ln(n, k, alpha, _ptr(A), max(1, n), beta, _ptr(C), max(1, n))
    return C


def dsyr2k(A, B, alpha=1.0, beta=0.0, C=None):
    A = _f64(A); B = _f64(B)
    n, k = A.shape
    C = np.zeros((n, n), order="F") if C is None else _f64(C).copy(order="F")
    lib().orc_dsyr2k_ln(n, k, alpha, _ptr(A), max(1, n), _ptr(B), max(1, n), beta, _ptr(C),
                        max(1, n))
    return C


def dtrmm(L, B, alpha=1.0):
    L = _f64(L); B = _f64(B)
    m, n = B.shape
    C = np.zeros((m, n), order="F")
    lib().orc_dtrmm_llnn(m, n, alpha, _ptr(L), max(1, m), _ptr(B), max(1, m), _ptr(C), max(1, m))
    return C


def dtrsm_rltn(L, B):
    """X with X*L^T = B (right, lower, transposed, non-unit)."""
    L = _f64(L)
    X = _f64(B).copy(order="F")
    m, n = X.shape
    lib().orc_dtrsm_rltn(m, n, _ptr(L), max(1, n), _ptr(X), max(1, m))
    return X


def dtrsm_llnn(L, B):
    """X with L*X = B (left, lower, not transposed, non-unit)."""
    L = _f64(L)
    X = _f64(B).copy(order="F")
    m, n = X.shape
    lib().orc_dtrsm_llnn(m, n, _ptr(L), max(1, m), _ptr(X), max(1, m))
    return X


def dsymm(A, B, alpha=1.0, beta=0.0, C=None):
    A = _f64(A); B = _f64(B)
    m, n = B.shape
    C = np.zeros((m, n), order="F") if C is None else _f64(C).copy(order="F")
    lib().orc_dsymm_ll(m, n, alpha, _ptr(A), max(1, m), _ptr(B), max(1, m), beta, _ptr(C),
                       max(1, m))
    return C


def tc_trsm_plain_emul(L, B, nb):
    """Plain-tier TRSM recursion (BLAS/tc_trsm.cu:31-137) with FP16-rounded updates."""
    L = _f64(L)
    X = _f64(B).copy(order="F")
    m, n = X.shape
    w = np.zeros(m * n + n * n)
    lib().orc_tc_trsm_plain_emul(m, n, _ptr(L), max(1, n), _ptr(X), max(1, m), nb, _ptr(w))
    return X


# ---------------------------------------------------------------------- Ozaki
def ozaki_slice(X, nslices, trans=False):
    """Slices of the rows of op(X): returns (slices[s, i, p] int8, exps[i] int32)."""
    X = _f64(X)
    m = X.shape[1] if trans else X.shape[0]
    k = X.shape[0] if trans else X.shape[1]
    sl = np.zeros((nslices, m, k), dtype=np.int8)
    ex = np.zeros(m, dtype=np.int32)
    lib().orc_ozaki_slice(int(trans), m, k, _ptr(X), max(1, X.shape[0]), nslices, _ptr(sl),
                          _ptr(ex))
    return sl, ex


def ozaki_gemm(sa, ea, sb, eb, alpha=1.0, beta=0.0, C=None, kchunk=0, ngroups=0):
    """C = alpha * (A_sliced * B_sliced^T) + beta*C from row slices of A (m x k) and
    row slices of B^T (n x k); slice-pair groups g < ngroups (default: nslices)."""
    ns, m, k = sa.shape
    n = sb.shape[1]
    sa = np.ascontiguousarray(sa); sb = np.ascontiguousarray(sb)
    ea = np.ascontiguousarray(ea, dtype=np.int32); eb = np.ascontiguousarray(eb, dtype=np.int32)
    C = np.zeros((m, n), order="F") if C is None else _f64(C).copy(order="F")
    lib().orc_ozaki_gemm(m, n, k, ns, _ptr(sa), _ptr(ea), _ptr(sb), _ptr(eb), alpha, beta,
                         _ptr(C), max(1, m), kchunk, int(ngroups))
    return C


# -------------------------------------------------------------------- metrics
def errors(X, R, S=None, lower=False):
    """(normwise, max-relative, componentwise) error of X against reference R."""
    X = _f64(X); R = _f64(R)
    m, n = R.shape
    out = np.zeros(3)
    Sp = None
    lds = 1
    if S is not None:
        S = _f64(S)
        Sp = _ptr(S)
        lds = max(1, S.shape[0])
    lib().orc_errors(m, n, _ptr(X), max(1, X.shape[0]), _ptr(R), max(1, R.shape[0]), Sp, lds,
                     int(lower), _ptr(out))
    return {"normwise": out[0], "maxrel": out[1], "componentwise": out[2]}


# ------------------------------------------------------------------ Cholesky
def dpotrf_lower(A):
    """(L, info): lower Cholesky factor of A (only A's lower triangle is read) in FP64;
    info = 0, or j+1 when the leading minor of order j+1 is not positive definite (L
    then holds the partial factorization).  L's strict upper triangle is zero."""
    L = np.tril(_f64(A)).copy(order="F")
    n = L.shape[0]
    info = lib().orc_dpotrf_lower(n, _ptr(L), max(1, n))
    return L, int(info)



# ------------------------------------------------- extended-precision columns
def rowdot_xp(X, Y, cols, X2=None, Y2=None):
    """out[:, c] = X @ Y[cols[c]] (+ X2 @ Y2[cols[c]]) with products and sums in long
    double (64-bit significand), rounded once to FP64.  X, Y, X2, Y2: 2-D float64
    arrays whose rows are the dot-product vectors (C order).  SYR2K's column j is
    rowdot_xp(A, B, [j], B, A) (n x k operands); column j of A @ B is
    rowdot_xp(A, B.T, [j])."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    n, k = X.shape
    assert Y.shape[1] == k
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    assert cols.min(initial=0) >= 0 and cols.max(initial=0) < Y.shape[0]
    if X2 is not None:
        X2 = np.ascontiguousarray(X2, dtype=np.float64)
        Y2 = np.ascontiguousarray(Y2, dtype=np.float64)
        assert X2.shape == X.shape and Y2.shape[1] == k and Y2.shape[0] == Y.shape[0]
    out = np.zeros((len(cols), n))
    lib().orc_rowdot_xp(n, k, _ptr(X), k, _ptr(Y), k, _ptr(X2) if X2 is not None else None, k,
                        _ptr(Y2) if Y2 is not None else None, k, len(cols), _ptr(cols), _ptr(out))
    return out.T.copy()
