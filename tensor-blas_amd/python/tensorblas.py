"""ctypes binding of the full C ABI of libTensorBLAS (include/TensorBLAS.h).

Pointers are device addresses (ints), e.g. ``tensor.data_ptr()`` of a torch tensor;
every matrix is column-major.  Every call raises ``TensorBLASError`` on a non-zero
status.  ``torch`` is imported before the library is loaded so that the process
keeps a single HIP runtime (torch's), which the library then shares.

This module is the host-side mirror of the reference's python/TensorBLAS.py for
all routines and tiers; ``TensorBLAS.py`` next to it keeps the reference's own
two functions unchanged.
"""
import ctypes
import os

import torch  # noqa: F401  (must be loaded first: one HIP runtime per process)

_HERE = os.path.dirname(os.path.realpath(__file__))
# the in-tree build; TENSORBLAS_LIB names another build of the same sources
LIB_PATH = os.environ.get("TENSORBLAS_LIB") or os.path.join(_HERE, "..", "build", "libTensorBLAS.so")

TIERS = {"plain": 0, "split": 1, "ozaki": 2, "fp32": 3, "fp64": 4}
F64_TIERS = {"ozaki", "fp64"}
STATUS = {0: "success", 1: "invalid argument", 2: "device error", 3: "out of device memory",
          4: "problem exceeds the no-allocation scratch"}


class TensorBLASError(RuntimeError):
    def __init__(self, fn, status):
        super().__init__(f"{fn} failed with status {status} ({STATUS.get(status, '?')})")
        self.status = status


_lib = None
_L, _D, _P, _I, _C = ctypes.c_long, ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_char


def lib():
    """Load libTensorBLAS.so (raises if it is not built: there is no fallback)."""
    global _lib
    if _lib is None:
        path = os.path.realpath(LIB_PATH)
        if not os.path.exists(path):
            raise RuntimeError(f"libTensorBLAS.so is not built ({path}); run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        sig = {
            "tb_create": [_P], "tb_destroy": [_P], "tb_set_stream": [_P, _P],
            "tb_synchronize": [_P], "tb_set_ozaki_slices": [_P, _I], "tb_set_ozaki_extra": [_P, _I],
            "tb_set_split_planes": [_P, _I],
            "tb_set_max_cus": [_P, _I],
            "tb_syrk": [_P, _I, _L, _L, _D, _P, _L, _D, _P, _L, _L],
            "tb_syrk_noalloc": [_P, _I, _L, _L, _D, _P, _L, _D, _P, _L, _L],
            "tb_syr2k": [_P, _I, _L, _L, _D, _P, _L, _P, _L, _D, _P, _L, _L],
            "tb_trmm": [_P, _I, _L, _L, _D, _P, _L, _P, _L, _P, _L, _L],
            "tb_trsm": [_P, _I, _L, _L, _P, _L, _P, _L, _L],
            "tb_trsm_refined": [_P, _I, _L, _L, _P, _L, _P, _L, _L],
            "tb_potrf": [_P, _I, _L, _P, _L, _L, ctypes.POINTER(ctypes.c_long)],
            "tb_trtri": [_P, _I, _C, _C, _L, _P, _L, _L, ctypes.POINTER(ctypes.c_long)],
            "tb_trsm_ex": [_P, _I] + [_C] * 5 + [_L, _L, _D, _P, _L, _P, _L, _L],
            "tb_trmm_ex": [_P, _I] + [_C] * 5 + [_L, _L, _D, _P, _L, _P, _L, _P, _L, _L],
            "tb_symm_ex": [_P, _I] + [_C] * 3 + [_L, _L, _D, _P, _L, _P, _L, _D, _P, _L],
            "tb_syrk_ex": [_P, _I] + [_C] * 3 + [_L, _L, _D, _P, _L, _D, _P, _L, _L],
            "tb_syr2k_ex": [_P, _I] + [_C] * 3 + [_L, _L, _D, _P, _L, _P, _L, _D, _P, _L, _L],
            "tb_symm": [_P, _I, _L, _L, _D, _P, _L, _P, _L, _D, _P, _L],
            "tb_gemm": [_P, _I, _I, _I, _L, _L, _L, _D, _P, _L, _P, _L, _D, _P, _L],
            "tb_s2h": [_P, _L, _L, _P, _L, _P, _L],
            "tb_s2h_transpose": [_P, _L, _L, _P, _P],
            "tb_copy_lower_to_upper": [_P, _I, _L, _P, _L],
            "tc_syrk_wrapper": [_L, _L, _P, _P, _L],
            "tc_trsm_wrapper": [_L, _L, _P, _P, _L],
            "tb_plan_chunks": [_I, _L, _P, _I],
            "tb_set_profiling": [_P, _I],
            "tb_profile_read": [_P, _P, _I],
            "tb_plan_trsm": [_L, _L, _P, _I],
            "tb_batch_create": [_P, _I, _P, _I, _L, _I, _P],
            "tb_batch_run": [_P],
            "tb_batch_info": [_P, _P, _I],
            "tb_batch_tiles": [_P, _P, _I],
            "tb_batch_destroy": [_P],
            "tb_serve": [_P, _I, _P, _P, _I, _P],
            "tb_reserve_workspace": [_P, _L, _L],
            "tb_workspace_sizes": [_P, _P],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _I
        L.tb_workspace_bytes.argtypes = [_P]
        L.tb_workspace_bytes.restype = _L
        L.tb_status_string.argtypes = [_I]
        L.tb_status_string.restype = ctypes.c_char_p
        L.tb_version.restype = _I
        L.print_env.restype = None
        _lib = L
    return _lib


def _c(ch):
    """one BLAS option character as a C char"""
    return ctypes.c_char(ch.encode() if isinstance(ch, str) else ch)


def _check(fn, st):
    if st != 0:
        raise TensorBLASError(fn, st)


def tier_id(tier):
    return TIERS[tier] if isinstance(tier, str) else int(tier)


class Handle:
    """A library handle (one HIP stream, one device workspace)."""

    def __init__(self, stream=None):
        self._h = ctypes.c_void_p()
        _check("tb_create", lib().tb_create(ctypes.byref(self._h)))
        if stream is not None:
            self.set_stream(stream)

    @property
    def ptr(self):
        return self._h

    def set_stream(self, stream):
        raw = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        _check("tb_set_stream", lib().tb_set_stream(self._h, raw))

    def synchronize(self):
        _check("tb_synchronize", lib().tb_synchronize(self._h))

    def set_ozaki_slices(self, s):
        _check("tb_set_ozaki_slices", lib().tb_set_ozaki_slices(self._h, s))

    def set_ozaki_extra(self, extra):
        """1: also sum the slice-pair group g = nslices (nslices 6..9); 0: the published
        triangular scheme."""
        _check("tb_set_ozaki_extra", lib().tb_set_ozaki_extra(self._h, int(extra)))

    def set_split_planes(self, n):
        _check("tb_set_split_planes", lib().tb_set_split_planes(self._h, n))

    def set_max_cus(self, n):
        """Cap the CUs this handle's GEMM kernels occupy (0: no cap); results do not
        change."""
        _check("tb_set_max_cus", lib().tb_set_max_cus(self._h, int(n)))

    PROF_NAMES = ["gemm_plain_sq", "gemm_plain_leaf", "gemm_split_sq", "gemm_split_leaf",
                  "gemm_ozaki_sq", "gemm_ozaki_leaf", "gemm_fp32_sq", "gemm_fp32_leaf",
                  "gemm_fp64_sq", "gemm_fp64_leaf", "ozaki_gemm", "ozaki_slice", "trtri", "planes", "batch",
                  "potrf_leaf"]

    def set_profiling(self, on):
        _check("tb_set_profiling", lib().tb_set_profiling(self._h, int(on)))

    def profile(self):
        """{category: (launches, total_ms, flops)} for categories that ran."""
        n = len(self.PROF_NAMES)
        out = (ctypes.c_double * (3 * n))()
        _check("tb_profile_read", lib().tb_profile_read(self._h, out, n))
        return {self.PROF_NAMES[c]: (int(out[3 * c]), out[3 * c + 1], out[3 * c + 2])
                for c in range(n) if out[3 * c] > 0}

    def workspace_bytes(self):
        return lib().tb_workspace_bytes(self._h)

    def close(self):
        if self._h:
            _check("tb_destroy", lib().tb_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ routines
    def syrk(self, tier, n, k, alpha, A, lda, beta, C, ldc, nb=256):
        _check("tb_syrk", lib().tb_syrk(self._h, tier_id(tier), n, k, alpha, A, lda, beta, C, ldc, nb))

    def syrk_noalloc(self, tier, n, k, alpha, A, lda, beta, C, ldc, nb=256):
        """tb_syrk without device allocation (fixed per-GPU scratch; see the header)"""
        _check("tb_syrk_noalloc", lib().tb_syrk_noalloc(self._h, tier_id(tier), n, k, alpha, A, lda,
                                                        beta, C, ldc, nb))

    def syr2k(self, tier, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb=256):
        _check("tb_syr2k", lib().tb_syr2k(self._h, tier_id(tier), n, k, alpha, A, lda, B, ldb, beta,
                                          C, ldc, nb))

    def trmm(self, tier, m, n, alpha, A, lda, B, ldb, C, ldc, nb=256):
        _check("tb_trmm", lib().tb_trmm(self._h, tier_id(tier), m, n, alpha, A, lda, B, ldb, C, ldc,
                                        nb))

    def trsm(self, tier, m, n, A, lda, B, ldb, nb=256):
        _check("tb_trsm", lib().tb_trsm(self._h, tier_id(tier), m, n, A, lda, B, ldb, nb))

    def trsm_refined(self, tier, m, n, A, lda, B, ldb, nb=256):
        _check("tb_trsm_refined", lib().tb_trsm_refined(self._h, tier_id(tier), m, n, A, lda, B, ldb,
                                                        nb))

    def potrf(self, tier, n, A, lda, nb=256):
        """Cholesky A = L L^T in place (lower); returns LAPACK's info (0: success, j: the
        leading minor of order j is not positive definite).  Synchronizes."""
        info = ctypes.c_long(-1)
        _check("tb_potrf", lib().tb_potrf(self._h, tier_id(tier), n, A, lda, nb, ctypes.byref(info)))
        return int(info.value)

    def trtri(self, tier, uplo, diag, n, A, lda, nb=256):
        """A := inv(A) for the uplo triangle; returns LAPACK's info (0, or i for a zero
        A(i, i), A unchanged)."""
        info = ctypes.c_long(0)
        _check("tb_trtri", lib().tb_trtri(self._h, tier_id(tier), _c(uplo), _c(diag), n, A, lda, nb,
                                          ctypes.byref(info)))
        return int(info.value)

    # BLAS option sets (characters as in BLAS: side L/R, uplo L/U, trans N/T, diag N/U;
    # order "C" column-major or "R" row-major as in CBLAS)
    def trsm_ex(self, tier, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, nb=256, order="C"):
        _check("tb_trsm_ex", lib().tb_trsm_ex(self._h, tier_id(tier), _c(order), _c(side), _c(uplo),
                                              _c(trans), _c(diag), m, n, alpha, A, lda, B, ldb, nb))

    def trmm_ex(self, tier, side, uplo, trans, diag, m, n, alpha, A, lda, B, ldb, C, ldc, nb=256,
                order="C"):
        _check("tb_trmm_ex", lib().tb_trmm_ex(self._h, tier_id(tier), _c(order), _c(side), _c(uplo),
                                              _c(trans), _c(diag), m, n, alpha, A, lda, B, ldb, C, ldc,
                                              nb))

    def symm_ex(self, tier, side, uplo, m, n, alpha, A, lda, B, ldb, beta, C, ldc, order="C"):
        _check("tb_symm_ex", lib().tb_symm_ex(self._h, tier_id(tier), _c(order), _c(side), _c(uplo), m,
                                              n, alpha, A, lda, B, ldb, beta, C, ldc))

    def syrk_ex(self, tier, uplo, trans, n, k, alpha, A, lda, beta, C, ldc, nb=256, order="C"):
        _check("tb_syrk_ex", lib().tb_syrk_ex(self._h, tier_id(tier), _c(order), _c(uplo), _c(trans), n,
                                              k, alpha, A, lda, beta, C, ldc, nb))

    def syr2k_ex(self, tier, uplo, trans, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb=256,
                 order="C"):
        _check("tb_syr2k_ex", lib().tb_syr2k_ex(self._h, tier_id(tier), _c(order), _c(uplo), _c(trans),
                                                n, k, alpha, A, lda, B, ldb, beta, C, ldc, nb))

    def symm(self, tier, m, n, alpha, A, lda, B, ldb, beta, C, ldc):
        _check("tb_symm", lib().tb_symm(self._h, tier_id(tier), m, n, alpha, A, lda, B, ldb, beta, C,
                                        ldc))

    def gemm(self, tier, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
        _check("tb_gemm", lib().tb_gemm(self._h, tier_id(tier), int(transa), int(transb), m, n, k,
                                        alpha, A, lda, B, ldb, beta, C, ldc))

    def s2h(self, m, n, As, ldas, Ah, ldah):
        _check("tb_s2h", lib().tb_s2h(self._h, m, n, As, ldas, Ah, ldah))

    def s2h_transpose(self, m, n, As, Ah):
        _check("tb_s2h_transpose", lib().tb_s2h_transpose(self._h, m, n, As, Ah))

    def reserve_workspace(self, nbytes, ozaki_bytes=0):
        _check("tb_reserve_workspace", lib().tb_reserve_workspace(self._h, nbytes, ozaki_bytes))

    def workspace_sizes(self):
        out = (ctypes.c_long * 2)()
        _check("tb_workspace_sizes", lib().tb_workspace_sizes(self._h, out))
        return int(out[0]), int(out[1])

    def copy_lower_to_upper(self, f64, n, A, lda):
        _check("tb_copy_lower_to_upper", lib().tb_copy_lower_to_upper(self._h, int(f64), n, A, lda))


# ------------------------------------------------------------------ batches
class BatchEntry(ctypes.Structure):
    """tb_batch_entry of include/TensorBLAS.h."""
    _fields_ = [("op", ctypes.c_int), ("m", ctypes.c_long), ("n", ctypes.c_long),
                ("k", ctypes.c_long), ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_long), ("C", ctypes.c_void_p),
                ("ldc", ctypes.c_long), ("B", ctypes.c_void_p), ("ldb", ctypes.c_long)]


BATCH_OPS = {"syrk": 0, "trsm": 1, "trmm": 2}
BATCH_AUTOTUNE, BATCH_NO_GRAPH = 1, 2
BATCH_TILES = {None: 0, "panel": 4, "square": 8}
BATCH_WAVES = {None: 0, 4: 16, 8: 32}


class Batch:
    """A batch plan: many SYRK / TRMM / TRSM problems run as one HIP graph (tb_batch_*).

    entries: dicts with op ('syrk' | 'trmm' | 'trsm'), n, and m (trmm, trsm) or k (syrk),
    A, lda, C, ldc, and B, ldb (trmm) as device addresses (for trsm A is L and C is B,
    for trmm A is L), alpha / beta (syrk; trmm alpha)."""

    def __init__(self, handle, tier, entries, nb=256, autotune=False, graph=True, tiles=None,
                 waves=None):
        arr = (BatchEntry * max(1, len(entries)))()
        for i, e in enumerate(entries):
            arr[i] = BatchEntry(BATCH_OPS[e["op"]], e.get("m", 0), e["n"], e.get("k", 0),
                                e.get("alpha", 1.0), e.get("beta", 0.0), e["A"], e["lda"], e["C"],
                                e["ldc"], e.get("B", 0) or None, e.get("ldb", 0))
        flags = ((BATCH_AUTOTUNE if autotune else 0) | (0 if graph else BATCH_NO_GRAPH) |
                 BATCH_TILES[tiles] | BATCH_WAVES[waves])
        self._b = ctypes.c_void_p()
        self._handle = handle          # the plan runs on the handle's stream
        _check("tb_batch_create", lib().tb_batch_create(handle.ptr, tier_id(tier), arr, len(entries),
                                                        nb, flags, ctypes.byref(self._b)))

    def run(self):
        _check("tb_batch_run", lib().tb_batch_run(self._b))

    def info(self):
        out = (ctypes.c_long * 6)()
        _check("tb_batch_info", lib().tb_batch_info(self._b, out, 6))
        return {"waves": out[0], "items": out[1], "instances": out[2], "square_tile_instances": out[3],
                "tune_us": out[4], "device_bytes": out[5]}

    def tiles(self):
        """the plan's tile choices: {'syrk': [shape per bucket], 'trmm': [...], 'waves': w};
        shape 0 = 64 x 256, 1 = 128 x 128; buckets n < 128, < 256, < 512, < 1024, < 2048, more"""
        out = (ctypes.c_int * 13)()
        _check("tb_batch_tiles", lib().tb_batch_tiles(self._b, out, 13))
        return {"syrk": list(out[0:6]), "trmm": list(out[6:12]), "waves": out[12]}

    def close(self):
        if self._b:
            _check("tb_batch_destroy", lib().tb_batch_destroy(self._b))
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- request loop
class Request(ctypes.Structure):
    """tb_request of include/TensorBLAS.h."""
    _fields_ = [("op", ctypes.c_int), ("tier", ctypes.c_int), ("m", ctypes.c_long),
                ("n", ctypes.c_long), ("k", ctypes.c_long), ("nb", ctypes.c_long),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("A", ctypes.c_void_p),
                ("lda", ctypes.c_long), ("B", ctypes.c_void_p), ("ldb", ctypes.c_long),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_long), ("arrival_us", ctypes.c_double),
                ("bytes", ctypes.c_long)]


class RequestResult(ctypes.Structure):
    _fields_ = [("admit_us", ctypes.c_double), ("done_us", ctypes.c_double),
                ("slot", ctypes.c_int), ("status", ctypes.c_int), ("est_us", ctypes.c_double)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_long), ("large_us", ctypes.c_double),
                ("max_large", ctypes.c_int), ("order", ctypes.c_int), ("reserve_cus", ctypes.c_int),
                ("partition", ctypes.c_int), ("small_us", ctypes.c_double),
                ("small_slots", ctypes.c_int)]


SERVE_ORDERS = {"fifo": 0, "shortest": 1, "shortest-small": 2}


REQ_OPS = {"syrk": 0, "syr2k": 1, "trmm": 2, "trsm": 3, "symm": 4}


def serve(slots, budget_bytes, requests, large_us=0.0, max_large=1, order="fifo", reserve_cus=0,
          partition=0, small_us=0.0, small_slots=0):
    """Run a request stream (tb_serve) on the given handles; requests are dicts with
    op, tier, m, n, k, nb, alpha, beta, A, lda, B, ldb, C, ldc, arrival_us, bytes.
    large_us > 0 puts the requests estimated above it in a large lane (at most max_large
    large requests at once); small_us > 0 those estimated below it in a small lane, for
    which small_slots slots are kept; order "fifo" admits the oldest waiting request of
    a lane first, "shortest" the one with the shortest estimated time alone,
    "shortest-small" the shortest in the small and middle lanes and the oldest in the
    large lane;
    reserve_cus > 0 keeps that many CUs free of large requests: of their GEMMs
    (partition 0, CU caps) or of all their kernels (partition 1, CU-masked streams);
    partition 2 keeps them free of the other requests instead.  Returns a list of dicts
    admit_us, done_us, slot, status, est_us."""
    n = len(requests)
    arr = (Request * max(1, n))()
    for i, r in enumerate(requests):
        arr[i] = Request(REQ_OPS[r["op"]], tier_id(r["tier"]), r.get("m", 0), r["n"], r.get("k", 0),
                         r.get("nb", 256), r.get("alpha", 1.0), r.get("beta", 0.0), r["A"],
                         r["lda"], r.get("B", 0) or 0, r.get("ldb", 1), r["C"], r["ldc"],
                         float(r.get("arrival_us", 0.0)), int(r.get("bytes", 0)))
    res = (RequestResult * max(1, n))()
    hs = (ctypes.c_void_p * len(slots))(*[s.ptr.value for s in slots])
    cfg = ServeConfig(int(budget_bytes), float(large_us), int(max_large), SERVE_ORDERS[order],
                      int(reserve_cus), int(partition), float(small_us), int(small_slots))
    _check("tb_serve", lib().tb_serve(hs, len(slots), ctypes.byref(cfg), arr, n, res))
    return [dict(admit_us=res[i].admit_us, done_us=res[i].done_us, slot=res[i].slot,
                 status=res[i].status, est_us=res[i].est_us) for i in range(n)]


# --------------------------------------------------------- host-only planners
def plan_chunks(n, trsm=False):
    """Chunks of n as the library schedules them (util/mat_size.cu:5-55 order)."""
    out = (ctypes.c_long * 64)()
    c = lib().tb_plan_chunks(int(trsm), n, out, 64)
    if c < 0:
        raise TensorBLASError("tb_plan_chunks", -c)
    return list(out[:c])


def plan_trsm(n, nb):
    """TRSM schedule: list of ('leaf', off, s) and ('update', c0, w, k0, kw) in order."""
    cap = 1 << 16
    out = (ctypes.c_long * (5 * cap))()
    c = lib().tb_plan_trsm(n, nb, out, cap)
    if c < 0:
        raise TensorBLASError("tb_plan_trsm", -c)
    ops = []
    for q in range(c):
        t, a, b, d, e = out[5 * q:5 * q + 5]
        ops.append(("leaf", a, b) if t == 0 else ("update", a, b, d, e))
    return ops
