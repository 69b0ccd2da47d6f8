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
