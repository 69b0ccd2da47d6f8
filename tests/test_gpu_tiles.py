"""Tile shapes and launch caps do not change results.

Launches of full-C plane products with at least 256 tiles of 256 x 128 run on those
wide tiles; smaller launches run on 128 x 128 tiles.  Every output element is summed
over k in the same order by either shape, so a product computed in one wide launch must
equal, bit for bit, the same product computed as column blocks small enough to stay on
square tiles.  The square-tile results are themselves held to the oracle's bounds in
test_gpu_blas3.py; here the wide result is also held to the same bound.  Likewise a
handle whose kernels are capped to a few CUs (tb_set_max_cus) computes the same bits.
"""
import numpy as np
import pytest

import oracle
from _tb import U32, gamma, lower_tri_dominant, rng, to_dev, from_dev, uniform

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def h():
    import tensorblas
    hd = tensorblas.Handle()
    yield hd
    hd.close()


# 4133 x 2057 outputs: 17 x 17 = 289 wide tiles (>= 256), so the whole product runs wide;
# a 128-column block is 17 wide tiles and runs on square tiles.
M, N, K = 4133, 2057, 200


@pytest.mark.parametrize("tier", ["plain", "split"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_gemm_wide_equals_square_blocks(h, tier, ta, tb):
    g = rng(4133 + 2 * ta + tb)
    A = uniform(g, (K, M) if ta else (M, K))
    B = uniform(g, (N, K) if tb else (K, N))
    C0 = uniform(g, (M, N))
    alpha, beta = 0.75, -1.25
    dA, lda = to_dev(A, A.shape[0] + 1)
    dB, ldb = to_dev(B, B.shape[0] + 3)
    dC, ldc = to_dev(C0, M + 2)
    dC2 = dC.clone()
    h.gemm(tier, ta, tb, M, N, K, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta,
           dC.data_ptr(), ldc)
    for j in range(0, N, 128):
        w = min(128, N - j)
        # columns j.. of op(B): column offset j of B (k x n), or row offset j of B (n x k)
        boff = (j if tb else j * ldb) * 4
        h.gemm(tier, ta, tb, M, w, K, alpha, dA.data_ptr(), lda, dB.data_ptr() + boff, ldb,
               beta, dC2.data_ptr() + j * ldc * 4, ldc)
    h.synchronize()
    C = from_dev(dC, M, N, ldc)
    C2 = from_dev(dC2, M, N, ldc)
    assert np.array_equal(C, C2), f"{np.sum(C != C2)} entries differ"
    # and the wide result against the FP64 oracle on a row sample
    opA = A.T if ta else A
    opB = B.T if tb else B
    rows = np.arange(0, M, 97)
    a = opA[rows].astype(np.float64)
    if tier == "plain":
        a = oracle.round_f16(opA[rows])
        b = oracle.round_f16(opB)
    else:
        b = opB.astype(np.float64)
    R = alpha * oracle.dgemm(a, b) + beta * C0[rows].astype(np.float64)
    S = abs(alpha) * oracle.dgemm(np.abs(a), np.abs(b)) + abs(beta) * np.abs(C0[rows])
    extra = 0.0
    if tier == "split":
        ra = np.abs(opA[rows]).max(axis=1)[:, None]
        cb = np.abs(opB).max(axis=0)[None, :]
        extra = abs(alpha) * 2.0 ** -21 * (ra * np.abs(opB).sum(axis=0)[None, :] +
                                           np.abs(opA[rows]).sum(axis=1)[:, None] * cb)
    err = np.abs(C[rows].astype(np.float64) - R)
    assert (err <= (gamma(K, U32) + 3 * U32) * S + extra).all()


@pytest.mark.parametrize("tier", ["plain", "split"])
def test_trsm_wide_updates_equal_row_blocks(h, tier):
    """X L^T = B at m = 8192, n = 4096: the top update (8192 x 2048, 512 wide tiles) runs
    wide; solving 1024-row blocks of B separately keeps every update on square tiles.
    Rows of X depend only on the same rows of B, so the results are equal bit for bit."""
    g = rng(8192 + (tier == "split"))
    m, n = 8192, 4096
    L = lower_tri_dominant(g, n)
    B = uniform(g, (m, n))
    dL, lda = to_dev(L)
    dB, ldb = to_dev(B)
    dB2 = dB.clone()
    h.trsm(tier, m, n, dL.data_ptr(), lda, dB.data_ptr(), ldb, nb=256)
    for r0 in range(0, m, 1024):
        h.trsm(tier, 1024, n, dL.data_ptr(), lda, dB2.data_ptr() + 4 * r0, ldb, nb=256)
    h.synchronize()
    X = from_dev(dB, m, n, ldb)
    X2 = from_dev(dB2, m, n, ldb)
    assert np.array_equal(X, X2), f"{np.sum(X != X2)} entries differ"
    rows = np.arange(0, m, 509)
    R = oracle.dtrsm_rltn(L.astype(np.float64), B[rows].astype(np.float64))
    e = oracle.errors(X[rows], R)
    tol = 2.0 ** -9 if tier == "plain" else 1e-6
    assert e["normwise"] < tol and e["maxrel"] < 4 * tol, e


@pytest.mark.parametrize("tier", ["plain", "split", "ozaki", "fp32", "fp64"])
def test_cu_cap_does_not_change_results(h, tier):
    """With tb_set_max_cus the GEMM kernels launch fewer workgroups than tiles and loop
    over them; every tile still sums k in the same order, so SYRK, TRSM and GEMM results
    equal the uncapped ones bit for bit."""
    import tensorblas
    g = rng(77 + len(tier))
    dt = np.float64 if tier in ("ozaki", "fp64") else np.float32
    n, k, m = 1500, 300, 700
    A = uniform(g, (n, k), dtype=dt)
    L = lower_tri_dominant(g, n, dt)
    B = uniform(g, (m, n), dtype=dt)
    outs = []
    for cap in (0, 24):
        hc = tensorblas.Handle()
        hc.set_max_cus(cap)
        dA, lda = to_dev(A)
        dC = torch.zeros(n * n, device="cuda", dtype=torch.float64 if dt == np.float64 else torch.float32)
        hc.syrk(tier, n, k, 1.0, dA.data_ptr(), lda, 0.0, dC.data_ptr(), n)
        dL, ldl = to_dev(L)
        dB, ldb = to_dev(B)
        hc.trsm(tier, m, n, dL.data_ptr(), ldl, dB.data_ptr(), ldb, nb=256)
        dD = torch.zeros_like(dC)
        hc.gemm(tier, 0, 1, n, n, k, 1.0, dA.data_ptr(), lda, dA.data_ptr(), lda, 0.0,
                dD.data_ptr(), n)
        hc.synchronize()
        outs.append((dC.cpu(), dB.cpu(), dD.cpu()))
        hc.close()
    for a, b, what in zip(outs[0], outs[1], ("syrk", "trsm", "gemm")):
        assert torch.equal(a, b), f"{tier} {what}: {int((a != b).sum())} entries differ"


@pytest.mark.parametrize("tier", ["plain", "split"])
def test_syrk_wide_lower_equals_square(h, tier):
    """SYRK and SYR2K at n = 5900: 24 x 25 / 2 = 300 lower-triangular wide tiles (>= 256),
    so the product runs on 256 x 256 tiles; a handle capped to some CUs keeps 128 x 128
    tiles.  Every element sums k in the same order: equal bit for bit, and the lower
    triangle only is written."""
    import tensorblas
    g = rng(5900 + (tier == "split"))
    n, k = 5900, 96
    A = uniform(g, (n, k))
    B = uniform(g, (n, k))
    outs = []
    for cap in (0, 64):
        hc = tensorblas.Handle()
        hc.set_max_cus(cap)
        dA, lda = to_dev(A)
        dB, ldb = to_dev(B)
        dC, ldc = to_dev(np.full((n, n), np.nan, np.float32))
        hc.syrk(tier, n, k, 1.0, dA.data_ptr(), lda, 0.0, dC.data_ptr(), ldc)
        dD, _ = to_dev(np.full((n, n), np.nan, np.float32))
        hc.syr2k(tier, n, k, 0.5, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.0, dD.data_ptr(), ldc)
        hc.synchronize()
        outs.append((from_dev(dC, n, n, ldc), from_dev(dD, n, n, ldc)))
        hc.close()
    for a, b, what in zip(outs[0], outs[1], ("syrk", "syr2k")):
        lo = np.tril_indices(n)
        assert np.array_equal(a[lo], b[lo]), f"{what}: {int(np.sum(a[lo] != b[lo]))} entries differ"
        assert np.isnan(np.triu(a, 1)[np.triu_indices(n, 1)]).all(), f"{what}: upper triangle written"
