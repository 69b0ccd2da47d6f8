"""Cholesky factorization (tb_potrf) against the oracle's FP64 Cholesky, per tier.

The factor of the (rounded) input is compared with the oracle's
(oracle.dpotrf_lower, pinned against numpy.linalg.cholesky in test_oracle.py) by the
relative Frobenius error of L; the strict upper triangle must come back untouched
(it holds NaN on input).  Tolerances: plain (FP16 products) 5e-3, split and fp32 2e-6,
fp64 1e-12, ozaki (7 slices by default) 1e-11; the inputs are well conditioned
(B B^T / n + I).
"""
import numpy as np
import pytest

import oracle
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


def spd(g, n, dt):
    B = g.standard_normal((n, n))
    return (B @ B.T / n + np.eye(n)).astype(dt)


def dtype_of(tier):
    return np.float64 if tier in ("fp64", "ozaki") else np.float32


@pytest.mark.parametrize("tier", ["plain", "split", "fp32", "fp64", "ozaki"])
@pytest.mark.parametrize("n", [1, 5, 128, 129, 300, 1000, 2053])
def test_potrf(h, tier, n):
    g = rng(n + 17)
    A = spd(g, n, dtype_of(tier))
    Ain = A.copy()
    Ain[np.triu_indices(n, 1)] = np.nan
    dA, lda = to_dev(Ain, n + 3)
    info = h.potrf(tier, n, dA.data_ptr(), lda)
    assert info == 0
    L = from_dev(dA, n, n, lda)
    assert np.isnan(L[np.triu_indices(n, 1)]).all()          # upper triangle untouched
    Lg = np.tril(L).astype(np.float64)
    R, rinfo = oracle.dpotrf_lower(A.astype(np.float64))
    assert rinfo == 0
    err = np.linalg.norm(Lg - R) / np.linalg.norm(R)
    assert err < TOL[tier], (tier, n, err)


def test_potrf_split_at_least_as_accurate_as_fp32(h):
    g = rng(4099)
    n = 1536
    A = spd(g, n, np.float32)
    R, _ = oracle.dpotrf_lower(A.astype(np.float64))
    errs = {}
    for tier in ("split", "fp32", "plain"):
        dA, lda = to_dev(A)
        assert h.potrf(tier, n, dA.data_ptr(), lda) == 0
        L = np.tril(from_dev(dA, n, n, lda)).astype(np.float64)
        errs[tier] = np.linalg.norm(L - R) / np.linalg.norm(R)
    assert errs["split"] <= 1.5 * errs["fp32"], errs
    assert errs["plain"] > errs["split"], errs


@pytest.mark.parametrize("tier", ["split", "fp32", "fp64", "ozaki", "plain"])
@pytest.mark.parametrize("n,k", [(60, 17), (300, 128), (300, 200), (700, 699)])
def test_potrf_info_first_non_pd_minor(h, tier, n, k):
    """A diagonal matrix with A(k, k) = -1: the leading minor of order k+1 is the first
    one not positive definite, wherever k falls (inside a diagonal block, on a block
    boundary, in the last block); the columns before k are factored."""
    A = np.eye(n, dtype=dtype_of(tier)) * 4.0
    A[k, k] = -1.0
    dA, lda = to_dev(A)
    assert h.potrf(tier, n, dA.data_ptr(), lda) == k + 1
    L = from_dev(dA, n, n, lda)
    assert np.array_equal(np.diag(L)[:k], np.full(k, 2.0, dtype=L.dtype))


def test_potrf_arguments(h):
    import tensorblas
    dA = torch.ones(4, device="cuda")
    assert h.potrf("split", 0, dA.data_ptr(), 1) == 0
    with pytest.raises(tensorblas.TensorBLASError):
        h.potrf("split", 4, dA.data_ptr(), 2)                  # lda < n
    with pytest.raises(tensorblas.TensorBLASError):
        h.potrf("split", -1, dA.data_ptr(), 1)
