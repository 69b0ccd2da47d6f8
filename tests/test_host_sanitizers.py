"""CPU test: the library's host sources built with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/cpp/Makefile target `san`, host code only) and run
without a GPU.  The driver (tests/cpp/host_san.cpp) plans the reference's schedules
(util/mat_size.cu:5-55, BLAS/tc_trsm.cu:31-137), checks every entry-point family's
argument checks and the no-device error paths; any sanitizer report fails the test,
and its planner output must equal the oracle's and the unsanitized library's."""
import os
import shutil
import subprocess

import pytest

import oracle
from _tb import LIB, ROOT

CPP = os.path.join(ROOT, "tests", "cpp")
EXE = os.path.join(CPP, "build", "host_san")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB) or shutil.which("make") is None
                                or not os.path.exists("/opt/rocm/bin/hipcc"),
                                reason="needs the built library and hipcc")


@pytest.fixture(scope="module")
def san_output():
    subprocess.run(["make", "-C", CPP, "san", "-j", "8"], check=True, stdout=subprocess.DEVNULL)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([EXE], capture_output=True, text=True, env=env, timeout=300)
    return p


def test_sanitized_host_code_runs_clean(san_output):
    p = san_output
    text = p.stdout + p.stderr
    assert p.returncode == 0, text[-4000:]
    for bad in ("AddressSanitizer", "LeakSanitizer", "runtime error:", "FAIL "):
        assert bad not in text, text[-4000:]
    assert "done 0" in p.stdout


def test_sanitized_planners_match_oracle_and_library(san_output):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tensor-blas_amd", "python"))
    import tensorblas

    chunks = trsm = 0
    for line in san_output.stdout.splitlines():
        if line.startswith("chunks "):
            head, _, tail = line.partition(":")
            _, planner, n = head.split()
            got = [int(v) for v in tail.split()]
            n = int(n)
            if planner == "0":
                assert got == tensorblas.plan_chunks(n), n
                if n <= 188416:            # the reference's power table
                    assert got == oracle.find_mat_size_syrk(n), n
            else:
                assert got == tensorblas.plan_chunks(n, trsm=True), n
                if n <= 161792 + 2048:     # the reference's TRSM power table
                    assert got == oracle.find_mat_size_trsm(n), n
            chunks += 1
        elif line.startswith("trsm "):
            head, _, tail = line.partition(":")
            _, n, nb = head.split()
            count, digest = (int(v) for v in tail.split())
            ops = tensorblas.plan_trsm(int(n), int(nb))
            s = 0
            for o in ops:
                vals = (0, o[1], o[2], 0, 0) if o[0] == "leaf" else (1,) + tuple(o[1:])
                for v in vals:
                    s = (s * 31 + v) % (1 << 64)
            assert (count, digest) == (len(ops), s), (n, nb)
            trsm += 1
    assert chunks == 28 and trsm == 8
