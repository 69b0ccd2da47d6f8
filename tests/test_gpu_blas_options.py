"""BLAS option sets (tb_*_ex: side, uplo, trans, diag) against scipy's FP64 BLAS.

Every option combination of TRSM, TRMM, SYMM, SYRK and SYR2K runs on the device and is
compared with scipy.linalg.blas (LAPACK's reference-interface dtrsm / dtrmm / dsymm /
dsyrk / dsyr2k) on the same inputs rounded to the tier's storage type, in column-major
and in row-major (CBLAS) storage.  The option
identities (transposition, reversal, unit diagonal) are exact, so the errors are those of
the native routine: relative Frobenius error within TOL[tier] (fp64 1e-12, Ozaki 1e-11,
split and fp32 2e-6, plain 5e-3).  The triangle a routine must not read holds NaN on input,
and the triangle of C that SYRK / SYR2K must not write must come back bit-identical.
"""
import itertools

import numpy as np
import pytest
from scipy.linalg import blas

from _tb import from_dev, rng, to_dev

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TIERS = ["split", "fp64", "ozaki", "plain"]
TOL = {"plain": 5e-3, "split": 2e-6, "fp32": 2e-6, "fp64": 1e-12, "ozaki": 1e-11}


@pytest.fixture(scope="module")
def h():
    import tensorblas
    hd = tensorblas.Handle()
    yield hd
    hd.close()


def dtype_of(tier):
    return np.float64 if tier in ("fp64", "ozaki") else np.float32


def rel(X, R):
    return np.linalg.norm(X - R) / max(np.linalg.norm(R), 1e-300)


def tri(g, n, lower, dt):
    """A well-conditioned triangle (also with a unit diagonal): off-diagonal U[-1,1]/n,
    diagonal 1 + U[0,1]; NaN in the other strict triangle."""
    T = g.uniform(-1, 1, (n, n)) / n + np.diag(1.0 + g.uniform(0, 1, n))
    T = (np.tril(T) if lower else np.triu(T)).astype(dt)
    T[np.triu_indices(n, 1) if lower else np.tril_indices(n, -1)] = np.nan
    return T


def clean(T):
    return np.nan_to_num(T).astype(np.float64)


def put(M, order, pad):
    """device copy of M in column-major ("C") or row-major ("R") storage, leading
    dimension = rows (resp. columns) + pad"""
    if order == "C":
        return to_dev(M, M.shape[0] + pad)
    return to_dev(M.T, M.shape[1] + pad)


def get(d, m, n, ld, order):
    return from_dev(d, m, n, ld) if order == "C" else from_dev(d, n, m, ld).T


OPTS4 = list(itertools.product("LR", "LU", "NT", "NU"))


@pytest.mark.parametrize("tier", TIERS)
@pytest.mark.parametrize("side,uplo,trans,diag", OPTS4)
@pytest.mark.parametrize("m,n", [(37, 70), (300, 45)])
@pytest.mark.parametrize("order", ["C", "R"])
def test_trsm_ex(h, order, tier, side, uplo, trans, diag, m, n):
    g = rng(m + 3 * n + ord(side) + 5 * ord(uplo) + 7 * ord(trans) + 11 * ord(diag))
    dt = dtype_of(tier)
    na = m if side == "L" else n
    A = tri(g, na, uplo == "L", dt)
    B = g.uniform(-1, 1, (m, n)).astype(dt)
    alpha = 0.75
    dA, lda = put(A, order, 2)
    dB, ldb = put(B, order, 1)
    h.trsm_ex(tier, side, uplo, trans, diag, m, n, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb,
              order=order)
    X = get(dB, m, n, ldb, order)
    R = blas.dtrsm(alpha, clean(A), B.astype(np.float64), side=side == "R", lower=uplo == "L",
                   trans_a=trans == "T", diag=diag == "U")
    assert rel(X, R) < TOL[tier], (tier, rel(X, R))


@pytest.mark.parametrize("tier", TIERS)
@pytest.mark.parametrize("side,uplo,trans,diag", OPTS4)
@pytest.mark.parametrize("m,n", [(37, 70), (300, 45)])
@pytest.mark.parametrize("order", ["C", "R"])
def test_trmm_ex(h, order, tier, side, uplo, trans, diag, m, n):
    g = rng(7 * m + n + ord(side) + 5 * ord(uplo) + 7 * ord(trans) + 11 * ord(diag))
    dt = dtype_of(tier)
    na = m if side == "L" else n
    A = tri(g, na, uplo == "L", dt)
    B = g.uniform(-1, 1, (m, n)).astype(dt)
    alpha = -1.25
    dA, lda = put(A, order, 1)
    dB, ldb = put(B, order, 2)
    dC, ldc = put(np.zeros((m, n), dt), order, 3)
    h.trmm_ex(tier, side, uplo, trans, diag, m, n, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb,
              dC.data_ptr(), ldc, order=order)
    C = get(dC, m, n, ldc, order)
    R = blas.dtrmm(alpha, clean(A), B.astype(np.float64), side=side == "R", lower=uplo == "L",
                   trans_a=trans == "T", diag=diag == "U")
    assert rel(C, R) < TOL[tier], (tier, rel(C, R))
    assert np.array_equal(get(dB, m, n, ldb, order), B)          # B is read only


@pytest.mark.parametrize("tier", TIERS)
@pytest.mark.parametrize("side,uplo", list(itertools.product("LR", "LU")))
@pytest.mark.parametrize("m,n,beta", [(37, 70, 0.5), (300, 45, 0.0)])
@pytest.mark.parametrize("order", ["C", "R"])
def test_symm_ex(h, order, tier, side, uplo, m, n, beta):
    g = rng(3 * m + n + ord(side) + 5 * ord(uplo))
    dt = dtype_of(tier)
    na = m if side == "L" else n
    S = g.uniform(-1, 1, (na, na))
    S = (S + S.T).astype(dt)
    A = S.copy()
    A[np.triu_indices(na, 1) if uplo == "L" else np.tril_indices(na, -1)] = np.nan
    B = g.uniform(-1, 1, (m, n)).astype(dt)
    C0 = g.uniform(-1, 1, (m, n)).astype(dt)
    dA, lda = put(A, order, 1)
    dB, ldb = put(B, order, 1)
    dC, ldc = put(C0 if beta else np.full((m, n), np.nan, dt), order, 2)
    h.symm_ex(tier, side, uplo, m, n, 1.5, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta,
              dC.data_ptr(), ldc, order=order)
    C = get(dC, m, n, ldc, order)
    R = blas.dsymm(1.5, S.astype(np.float64), B.astype(np.float64), beta, C0.astype(np.float64),
                   side=side == "R", lower=uplo == "L")
    assert rel(C, R) < TOL[tier], (tier, rel(C, R))


def sym_case(g, n, uplo, dt):
    """C0 with NaN in the triangle that must stay untouched (and not be read)."""
    C0 = g.uniform(-1, 1, (n, n)).astype(dt)
    off = np.triu_indices(n, 1) if uplo == "L" else np.tril_indices(n, -1)
    C0[off] = np.nan
    return C0, off


@pytest.mark.parametrize("tier", TIERS)
@pytest.mark.parametrize("uplo,trans", list(itertools.product("LU", "NT")))
@pytest.mark.parametrize("n,k", [(45, 70), (300, 33)])
@pytest.mark.parametrize("order", ["C", "R"])
def test_syrk_ex(h, order, tier, uplo, trans, n, k):
    g = rng(n + 5 * k + ord(uplo) + 3 * ord(trans))
    dt = dtype_of(tier)
    A = g.uniform(-1, 1, (n, k) if trans == "N" else (k, n)).astype(dt)
    C0, off = sym_case(g, n, uplo, dt)
    dA, lda = put(A, order, 1)
    dC, ldc = put(C0, order, 2)
    h.syrk_ex(tier, uplo, trans, n, k, 0.5, dA.data_ptr(), lda, 2.0, dC.data_ptr(), ldc, order=order)
    C = get(dC, n, n, ldc, order)
    R = blas.dsyrk(0.5, A.astype(np.float64), 2.0, np.nan_to_num(C0).astype(np.float64),
                   trans=int(trans == "T"), lower=int(uplo == "L"))
    keep = np.tril if uplo == "L" else np.triu
    assert rel(keep(C), keep(R)) < TOL[tier], (tier, rel(keep(C), keep(R)))
    assert np.isnan(C[off]).all()


@pytest.mark.parametrize("tier", TIERS)
@pytest.mark.parametrize("uplo,trans", list(itertools.product("LU", "NT")))
@pytest.mark.parametrize("n,k", [(45, 70), (300, 33)])
@pytest.mark.parametrize("order", ["C", "R"])
def test_syr2k_ex(h, order, tier, uplo, trans, n, k):
    g = rng(3 * n + k + ord(uplo) + 5 * ord(trans))
    dt = dtype_of(tier)
    shape = (n, k) if trans == "N" else (k, n)
    A = g.uniform(-1, 1, shape).astype(dt)
    B = g.uniform(-1, 1, shape).astype(dt)
    C0, off = sym_case(g, n, uplo, dt)
    dA, lda = put(A, order, 1)
    dB, ldb = put(B, order, 3)
    dC, ldc = put(C0, order, 1)
    h.syr2k_ex(tier, uplo, trans, n, k, -0.5, dA.data_ptr(), lda, dB.data_ptr(), ldb, 1.0,
               dC.data_ptr(), ldc, order=order)
    C = get(dC, n, n, ldc, order)
    R = blas.dsyr2k(-0.5, A.astype(np.float64), B.astype(np.float64), 1.0,
                    np.nan_to_num(C0).astype(np.float64), trans=int(trans == "T"),
                    lower=int(uplo == "L"))
    keep = np.tril if uplo == "L" else np.triu
    assert rel(keep(C), keep(R)) < TOL[tier], (tier, rel(keep(C), keep(R)))
    assert np.isnan(C[off]).all()


def test_native_options_equal_native_routine(h):
    """(side R, uplo L, trans T, diag N) is the native TRSM: bit-identical results."""
    g = rng(99)
    m, n = 40, 300
    A = tri(g, n, True, np.float32)
    B = g.uniform(-1, 1, (m, n)).astype(np.float32)
    dA, lda = to_dev(A)
    d1, ldb = to_dev(B)
    d2, _ = to_dev(B)
    h.trsm("split", m, n, dA.data_ptr(), lda, d1.data_ptr(), ldb)
    h.trsm_ex("split", "R", "L", "T", "N", m, n, 1.0, dA.data_ptr(), lda, d2.data_ptr(), ldb)
    assert torch.equal(d1.nan_to_num(7.0), d2.nan_to_num(7.0))


def test_options_lowercase_conjugate_and_alpha_zero(h):
    g = rng(5)
    m, n = 20, 30
    A = tri(g, n, False, np.float64)
    B = g.uniform(-1, 1, (m, n))
    dA, lda = to_dev(A)
    dB, ldb = to_dev(B)
    h.trsm_ex("fp64", "r", "u", "c", "n", m, n, 2.0, dA.data_ptr(), lda, dB.data_ptr(), ldb)
    R = blas.dtrsm(2.0, clean(A), B, side=1, lower=0, trans_a=1)
    assert rel(from_dev(dB, m, n, ldb), R) < 1e-12
    h.trsm_ex("fp64", "L", "L", "N", "N", m, n, 0.0, 0, m, dB.data_ptr(), ldb)  # A not read
    assert not from_dev(dB, m, n, ldb).any()


def test_options_arguments(h):
    import tensorblas
    d = torch.zeros(64, dtype=torch.float64, device="cuda")
    p = d.data_ptr()
    bad = [
        lambda: h.trsm_ex("fp64", "X", "L", "N", "N", 4, 4, 1.0, p, 4, p, 4),
        lambda: h.trsm_ex("fp64", "L", "L", "N", "N", 4, 4, 1.0, p, 3, p, 4),   # lda < m
        lambda: h.trsm_ex("fp64", "R", "L", "N", "N", 2, 4, 1.0, p, 3, p, 2),   # lda < n
        lambda: h.trmm_ex("fp64", "L", "Q", "N", "N", 4, 4, 1.0, p, 4, p, 4, p, 4),
        lambda: h.symm_ex("fp64", "L", "L", 4, 4, 1.0, p, 4, p, 3, 0.0, p, 4),  # ldb < m
        lambda: h.syrk_ex("fp64", "U", "T", 4, 8, 1.0, p, 4, 0.0, p, 4),        # lda < k
        lambda: h.syr2k_ex("fp64", "L", "X", 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4),
        lambda: h.syrk_ex("fp64", "L", "N", 4, 4, 1.0, p, 4, 0.0, p, 4, order="X"),
        lambda: h.trsm_ex("fp64", "L", "L", "N", "N", 4, 2, 1.0, p, 2, p, 2, order="R"),  # lda < m
        lambda: h.trsm_ex("fp64", "L", "L", "N", "N", 4, 2, 1.0, p, 4, p, 1, order="R"),  # ldb < n
    ]
    for f in bad:
        with pytest.raises(tensorblas.TensorBLASError):
            f()
    h.syrk_ex("fp64", "U", "N", 0, 4, 1.0, p, 1, 0.0, p, 1)                   # empty: no-op


@pytest.mark.parametrize("tier", ["split", "fp64"])
def test_null_operands_rejected_before_any_copy(h, tier):
    """Option sets that copy operands (symm side R or uplo U, syrk uplo U, syr2k trans T)
    reject a null operand with TB_ERR_INVALID before the copy is launched; the caller's
    other buffers are untouched."""
    import tensorblas
    L = tensorblas.lib()
    c = tensorblas._c
    t = tensorblas.tier_id(tier)
    dt = torch.float64 if tier == "fp64" else torch.float32
    d = torch.full((64,), 7.0, dtype=dt, device="cuda")
    p = d.data_ptr()
    hp = h.ptr
    calls = [
        lambda: L.tb_symm_ex(hp, t, c("C"), c("R"), c("L"), 4, 4, 1.0, None, 4, p, 4, 0.0, p, 4),
        lambda: L.tb_symm_ex(hp, t, c("C"), c("L"), c("U"), 4, 4, 1.0, p, 4, None, 4, 0.0, p, 4),
        lambda: L.tb_symm_ex(hp, t, c("C"), c("R"), c("U"), 4, 4, 1.0, p, 4, p, 4, 0.0, None, 4),
        lambda: L.tb_syrk_ex(hp, t, c("C"), c("U"), c("N"), 4, 4, 1.0, p, 4, 0.0, None, 4, 256),
        lambda: L.tb_syrk_ex(hp, t, c("C"), c("U"), c("T"), 4, 4, 1.0, None, 4, 0.0, p, 4, 256),
        lambda: L.tb_syr2k_ex(hp, t, c("C"), c("L"), c("T"), 4, 4, 1.0, None, 4, p, 4, 0.0, p, 4, 256),
        lambda: L.tb_syr2k_ex(hp, t, c("C"), c("U"), c("T"), 4, 4, 1.0, p, 4, None, 4, 0.0, p, 4, 256),
        lambda: L.tb_syr2k_ex(hp, t, c("C"), c("U"), c("N"), 4, 4, 1.0, p, 4, p, 4, 0.0, None, 4, 256),
    ]
    for f in calls:
        assert f() == 1
    h.synchronize()
    assert bool((d == 7.0).all())
