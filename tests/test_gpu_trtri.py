"""Triangular inverse (tb_trtri) against the FP64 inverse numpy (LAPACK) forms from the
same rounded input, per tier, both triangles, non-unit and unit diagonals.

The inverse is unique, so the comparison pins the result; tolerances on the relative
Frobenius error: plain 5e-3, split and fp32 2e-6, fp64 1e-12, ozaki 1e-11 (the inputs
are well conditioned: off-diagonal U[-1,1]/n, diagonal 1 + U[0,1]).  The other strict
triangle holds NaN and must come back NaN; with a unit diagonal the stored diagonal
holds 7 and must come back 7.
"""
import numpy as np
import pytest

from _tb import from_dev, rng, to_dev

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = {"plain": 5e-3, "split": 2e-6, "fp32": 2e-6, "fp64": 1e-12, "ozaki": 1e-11}


@pytest.fixture(scope="module")
def h():
    import tensorblas
    hd = tensorblas.Handle()
    yield hd
    hd.close()


def dtype_of(tier):
    return np.float64 if tier in ("fp64", "ozaki") else np.float32


def case(g, n, lower, unit, dt):
    T = g.uniform(-1, 1, (n, n)) / n + np.diag(1.0 + g.uniform(0, 1, n))
    T = (np.tril(T) if lower else np.triu(T)).astype(dt)
    Tm = T.astype(np.float64)
    if unit:
        np.fill_diagonal(Tm, 1.0)
        np.fill_diagonal(T, 7.0)
    other = np.triu_indices(n, 1) if lower else np.tril_indices(n, -1)
    T[other] = np.nan
    Tm[other] = 0.0
    return T, Tm, other


@pytest.mark.parametrize("tier", ["split", "fp64", "ozaki", "plain", "fp32"])
@pytest.mark.parametrize("uplo,diag", [("L", "N"), ("U", "N"), ("L", "U"), ("U", "U")])
@pytest.mark.parametrize("n", [1, 5, 256, 257, 300, 1000])
def test_trtri(h, tier, uplo, diag, n):
    g = rng(n * 3 + ord(uplo) + 5 * ord(diag))
    T, Tm, other = case(g, n, uplo == "L", diag == "U", dtype_of(tier))
    dA, lda = to_dev(T, n + 5)
    assert h.trtri(tier, uplo, diag, n, dA.data_ptr(), lda) == 0
    X = from_dev(dA, n, n, lda)
    assert np.isnan(X[other]).all()
    if diag == "U":
        assert (np.diag(X) == 7.0).all()
        np.fill_diagonal(X, 1.0)
    X = np.nan_to_num(X).astype(np.float64)
    R = np.linalg.inv(Tm)
    err = np.linalg.norm(X - R) / np.linalg.norm(R)
    assert err < TOL[tier], (tier, uplo, diag, n, err)


@pytest.mark.parametrize("tier", ["split", "ozaki"])
@pytest.mark.parametrize("k", [0, 77, 256, 599])
def test_trtri_singular(h, tier, k):
    n = 600
    g = rng(k)
    T, _, _ = case(g, n, True, False, dtype_of(tier))
    T[k, k] = 0.0
    if k + 10 < n:
        T[k + 10, k + 10] = 0.0                 # a later zero: the first one is reported
    dA, lda = to_dev(T)
    before = dA.clone()
    assert h.trtri(tier, "L", "N", n, dA.data_ptr(), lda) == k + 1
    assert torch.equal(before.nan_to_num(3.0), dA.nan_to_num(3.0))    # A unchanged


def test_trtri_tensor_api():
    import TensorBLAS as TB
    g = rng(8)
    n = 300
    for upper in (False, True):
        for col in (False, True):
            T, Tm, _ = case(g, n, not upper, False, np.float64)
            A = torch.from_numpy(np.nan_to_num(T)).cuda()
            if col:
                A = A.T.contiguous().T
            assert TB.trtri(A, upper=upper, tier="fp64") == 0
            X = (torch.triu(A) if upper else torch.tril(A)).cpu().numpy()
            R = np.linalg.inv(Tm)
            assert np.linalg.norm(X - R) / np.linalg.norm(R) < 1e-12, (upper, col)


def test_trtri_arguments(h):
    import tensorblas
    d = torch.ones(16, dtype=torch.float64, device="cuda")
    assert h.trtri("fp64", "L", "N", 0, d.data_ptr(), 1) == 0
    for args in [("fp64", "X", "N", 4, d.data_ptr(), 4), ("fp64", "L", "Q", 4, d.data_ptr(), 4),
                 ("fp64", "L", "N", 4, d.data_ptr(), 3), ("fp64", "L", "N", -1, d.data_ptr(), 1)]:
        with pytest.raises(tensorblas.TensorBLASError):
            h.trtri(*args)
