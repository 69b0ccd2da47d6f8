"""CPU checks of the torch-tensor binding (tensor-blas_amd/python/TensorBLAS.py): every
call validates its tensors before anything reaches the library, so these run without a
GPU; the computations themselves are in test_gpu_torch_api.py."""
import os

import pytest

from _tb import LIB

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libTensorBLAS.so not built")


@pytest.fixture(scope="module")
def TB():
    import TensorBLAS
    return TensorBLAS


def test_exports_every_routine(TB):
    for name in ("tc_syrk", "tc_trsm", "tc_syr2k", "tc_trmm", "tc_symm", "syrk", "syr2k", "trmm",
                 "trsm", "symm", "potrf"):
        assert callable(getattr(TB, name)), name


@pytest.mark.parametrize("call", [
    lambda TB, A: TB.syrk(A),                       # CPU tensor
    lambda TB, A: TB.syr2k(A, A),
    lambda TB, A: TB.trmm(A, A),
    lambda TB, A: TB.trsm(A, A),
    lambda TB, A: TB.symm(A, A),
    lambda TB, A: TB.potrf(A),
    lambda TB, A: TB.tc_syr2k(A, A, tier="split"),
    lambda TB, A: TB.tc_trmm(A, A, tier="fp64"),    # wrong dtype as well
    lambda TB, A: TB.syrk(A, tier="bogus"),
    lambda TB, A: TB.syrk(A[0]),                    # 1-D
])
def test_rejected_before_the_library(TB, call):
    with pytest.raises(ValueError):
        call(TB, torch.ones(4, 4))


def test_layout_detection(TB):
    t = torch.zeros(5, 7)
    assert TB._layout(t) == ("R", 7)
    assert TB._layout(t.T) == ("C", 7)
    assert TB._layout(torch.zeros(8, 9)[:5, :7]) == ("R", 9)
    assert TB._layout(torch.zeros(8, 8)[::2, ::2]) is None
    a, ld = TB._as(t, "C")
    assert TB._layout(a) == ("C", 5) and ld == 5 and torch.equal(a, t)


class _Recorder:
    """stands in for tensorblas.Handle: records the C-ABI calls the binding makes"""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def rec(*args, **kw):
            self.calls.append((name, args, kw))
            return 0
        return rec


@pytest.fixture
def rec(TB, monkeypatch):
    r = _Recorder()
    monkeypatch.setattr(TB, "_handle", lambda t: r)
    monkeypatch.setattr(TB, "_check", lambda tier, *ts: None)    # CPU tensors pass
    return r


def test_row_major_calls_pass_order_R_and_row_strides(TB, rec):
    A = torch.zeros(6, 6)
    B = torch.zeros(6, 9)
    TB.trsm(A, B, side="L", uplo="U", trans="T", tier="split")
    name, args, kw = rec.calls[-1]
    assert name == "trsm_ex" and kw["order"] == "R"
    assert args[1:5] == ("L", "U", "T", "N") and args[5:7] == (6, 9)
    assert args[9] == 6 and args[11] == 9                         # lda, ldb: row strides


def test_output_order_decides_and_inputs_are_copied(TB, rec):
    A = torch.zeros(5, 5)                                         # row-major input
    B = torch.zeros(7, 5).T                                       # column-major 5 x 7 output
    TB.trsm(A, B, side="L", tier="split")
    name, args, kw = rec.calls[-1]
    assert kw["order"] == "C"
    assert args[9] == 5 and args[11] == 5                         # A copied column-major


def test_potrf_and_trtri_triangles(TB, rec):
    # column-major memory of a row-major tensor is its transpose: a row-major upper
    # triangle is the library's (column-major) lower one
    A = torch.zeros(4, 4)
    TB.potrf(A, upper=True, tier="split")
    assert rec.calls[-1][0] == "potrf"
    TB.trtri(A, upper=False, tier="split")
    assert rec.calls[-1][0] == "trtri" and rec.calls[-1][1][1] == "U"
    TB.trtri(A.T, upper=False, tier="split")                      # column-major view
    assert rec.calls[-1][1][1] == "L"
    n_before = len(rec.calls)
    TB.potrf(A, upper=False, tier="split")                        # the other triangle:
    assert rec.calls[n_before][0] == "potrf"                      # factors a column-major copy
    assert rec.calls[n_before][1][3] == 4


def test_syrk_shapes_follow_trans(TB, rec):
    A = torch.zeros(3, 8)                                         # k x n for trans T
    C = TB.syrk(A, trans="T", tier="split")
    assert tuple(C.shape) == (8, 8)
    name, args, kw = rec.calls[-1]
    assert name == "syrk_ex" and args[3:5] == (8, 3) and kw["order"] == "R"
