"""CPU tests of the C-ABI library: it loads, exports what include/TensorBLAS.h
declares, rejects bad arguments before touching a device, and plans the reference's
schedules (util/mat_size.cu:5-55, BLAS/tc_trsm.cu:31-137)."""
import ctypes
import re
import subprocess

import numpy as np
import pytest

import oracle
from _tb import HEADER, LIB

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(LIB),
                                reason="libTensorBLAS.so not built")


def header_functions():
    txt = open(HEADER).read()
    c_part = txt.split("#ifdef __cplusplus\n} /* extern \"C\" */")[0]
    cpp_part = txt.split("#ifdef __cplusplus\n} /* extern \"C\" */")[1]
    c_names = re.findall(r"TB_API\s+[\w\s\*]+?\b(\w+)\s*\(", c_part)
    cpp_names = re.findall(r"TB_API\s+int\s+(\w+)\s*\(", cpp_part)
    return c_names, cpp_names


def test_header_declares_reference_entry_points():
    c_names, cpp_names = header_functions()
    # include/TensorBLAS.h:135-167 of the reference
    for f in ("print_env", "tc_syrk_wrapper", "tc_trsm_wrapper"):
        assert f in c_names
    ref_cpp = ["tc_syrk", "tc_cumpsgemm_syrk", "tc_ozimmu_syrk", "tc_trsm", "tc_cumpsgemm_trsm",
               "tc_ozimmu_trsm", "tc_trmm", "tc_cumpsgemm_trmm", "tc_ozimmu_trmm", "tc_syr2k",
               "tc_cumpsgemm_syr2k", "tc_ozimmu_syr2k", "tc_symm", "tc_cumpsgemm_symm"]
    for f in ref_cpp + ["tc_ozimmu_symm"]:
        assert f in cpp_names


def test_library_exports_every_declared_symbol():
    c_names, cpp_names = header_functions()
    L = ctypes.CDLL(LIB)
    for f in c_names:
        assert hasattr(L, f), f
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split(" ", 2)[2].split("(")[0] for line in out.splitlines()
                if " T " in line}
    for f in cpp_names:
        assert f in exported, f
    # nothing internal leaks out
    assert not any(e.startswith("tb::") for e in exported)


def test_tensorblas_so_name_python_loads():
    # python/TensorBLAS.py:9 loads ../build/TensorBLAS.so (the Makefile built libTensorBLAS.so)
    import os
    alt = os.path.join(os.path.dirname(LIB), "TensorBLAS.so")
    assert os.path.exists(alt)
    assert hasattr(ctypes.CDLL(alt), "tc_syrk_wrapper")


def _lib():
    import tensorblas
    return tensorblas


def test_invalid_arguments_rejected_without_device():
    tbm = _lib()
    L = tbm.lib()
    null = ctypes.c_void_p()
    assert L.tb_syrk(null, 0, 8, 8, 1.0, 1, 8, 0.0, 1, 8, 256) == 1
    assert L.tb_trsm(null, 1, 8, 8, 1, 8, 1, 8, 256) == 1
    assert L.tb_set_ozaki_slices(null, 7) == 1
    assert L.tb_status_string(1) == b"invalid argument"
    assert L.tb_version() >= 100
    # the option-set, factorization and inverse entry points check the handle first
    c = ctypes.c_char
    assert L.tb_trsm_ex(null, 1, c(b"C"), c(b"L"), c(b"L"), c(b"N"), c(b"N"), 8, 8, 1.0, 1, 8, 1, 8, 256) == 1
    assert L.tb_syrk_ex(null, 1, c(b"R"), c(b"U"), c(b"T"), 8, 8, 1.0, 1, 8, 0.0, 1, 8, 256) == 1
    info = ctypes.c_long(0)
    assert L.tb_potrf(null, 1, 8, 1, 8, 256, ctypes.byref(info)) == 1
    assert L.tb_trtri(null, 1, c(b"L"), c(b"N"), 8, 1, 8, 256, ctypes.byref(info)) == 1


@pytest.mark.parametrize("std,lang,compiler", [("c99", "c", "gcc"), ("c++17", "c++", "g++")])
def test_header_compiles_standalone(tmp_path, std, lang, compiler):
    """include/TensorBLAS.h is the boundary a C, cgo or JNI caller compiles against:
    it must build on its own, warning-free, as C99 and as C++."""
    import os
    import shutil
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not installed")
    src = tmp_path / ("h." + ("c" if lang == "c" else "cpp"))
    src.write_text('#include "TensorBLAS.h"\nint main(void) { return TB_SUCCESS; }\n')
    r = subprocess.run([compiler, f"-std={std}", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.dirname(HEADER), "-c", str(src), "-o", str(tmp_path / "h.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("n", [0, 1, 2048, 8191, 8192, 16384, 20000, 32768, 40960, 188416])
def test_plan_chunks_syrk_matches_oracle(n):
    assert _lib().plan_chunks(n) == oracle.find_mat_size_syrk(n)


@pytest.mark.parametrize("n", [0, 300, 2048, 6144, 16384, 20000, 30720, 161792])
def test_plan_chunks_trsm_matches_oracle(n):
    assert _lib().plan_chunks(n, trsm=True) == oracle.find_mat_size_trsm(n)


def test_plan_chunks_beyond_reference_table():
    # the reference reads powers[-1] there; the library keeps taking 16384s
    assert _lib().plan_chunks(188416 + 8192 + 5) == [16384] * 11 + [8192, 8197]
    assert _lib().plan_chunks(188416 + 20000) == [16384] * 11 + [8192, 16384, 3616]
    assert _lib().plan_chunks(161792 + 2048, trsm=True) == [16384] * 9 + [8192, 4096, 2048, 2048]


def _check_trsm_schedule(n, nb, ops):
    """Every column solved once by a leaf, in order; every update reads only solved
    columns and writes only unsolved ones; each (target, source) column pair of the
    strict lower triangle of L is applied exactly once."""
    solved = np.zeros(n, dtype=bool)
    applied = np.zeros((n, n), dtype=np.int32)
    for op in ops:
        if op[0] == "leaf":
            _, off, s = op
            assert 0 < s <= max(nb, 1) or s <= 256
            assert not solved[off:off + s].any()
            solved[off:off + s] = True
            assert applied[off:off + s, :off].sum() == s * off   # all earlier sources applied
            # the leaf itself resolves the couplings inside its diagonal block
            applied[off:off + s, off:off + s] += np.tril(np.ones((s, s), dtype=np.int32), -1)
        else:
            _, c0, w, k0, kw = op
            assert solved[k0:k0 + kw].all() and not solved[c0:c0 + w].any()
            applied[c0:c0 + w, k0:k0 + kw] += 1
    assert solved.all()
    assert np.array_equal(applied, np.tril(np.ones((n, n), dtype=np.int32), -1))


@pytest.mark.parametrize("n,nb", [(16384, 256), (2048, 256), (20000, 256), (6144, 512),
                                  (300, 256), (4096, 128), (2048, 2048)])
def test_trsm_schedule(n, nb):
    ops = _lib().plan_trsm(n, nb)
    leaves = [o for o in ops if o[0] == "leaf"]
    ups = [o for o in ops if o[0] == "update"]
    if n <= 6144:
        _check_trsm_schedule(n, nb, ops)
    if (n, nb) == (16384, 256):
        # BLAS/tc_trsm.cu:31-80 at the TRSM benchmark size: 64 leaves, 63 update GEMMs
        assert len(leaves) == 64 and len(ups) == 63
        assert all(o[2] == 256 for o in leaves)
        assert ups[-1] == ("update", 16128, 256, 15872, 256)
        assert ("update", 8192, 8192, 0, 8192) in ups
    if (n, nb) == (20000, 256):
        # chunks processed smallest first: 1568 (one FP32 leaf block), 2048, 16384
        assert leaves[0][1] == 0
        assert ("update", 1568, 18432, 0, 1568) in ups
        assert ("update", 3616, 16384, 1568, 2048) in ups


def test_trsm_schedule_leaves_capped_at_256():
    ops = _lib().plan_trsm(4096, 4096)
    assert max(o[2] for o in ops if o[0] == "leaf") == 256
