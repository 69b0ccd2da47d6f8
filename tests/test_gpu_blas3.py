"""GPU parity of every routine and tier against the CPU oracle, through the C ABI.

Bounds (componentwise, S = |alpha|*|op(A)||op(B)| + |beta|*|C0| in FP64):
  plain  : FP64 oracle on the FP16-rounded operands (what the reference's s2h +
           FP16 GEMM computes); |C - C16| <= (gamma_k(2^-24) + 3*2^-24) * S16.
           This isolates the FP32 accumulation; a wrong index or transpose is O(1).
  split  : FP64 oracle on the FP32 operands; each operand is hi + 2^-11 lo with
           relative error <= 2^-22 and the dropped lo*lo term is <= 2^-22, so
           |C - C64| <= (2^-20 + gamma_k(2^-24) + 3*2^-24) * S.
  fp32   : |C - C64| <= (gamma_k(2^-24) + 3*2^-24) * S.
  fp64   : |C - C64| <= (gamma_k(2^-53) + 3*2^-53) * S.
  ozaki  : bit-exact equality with oracle.ozaki_gemm on the oracle's own slices.
"""
import numpy as np
import pytest

import oracle
from _tb import U32, U64, from_dev, gamma, lower_tri_dominant, rng, spread, to_dev, uniform

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def tbm():
    import tensorblas
    return tensorblas


@pytest.fixture(scope="module")
def h(tbm):
    hd = tbm.Handle()
    yield hd
    hd.close()


def _ozaki_kchunk(S):
    return ((2 ** 31 - 1) // (S * 127 * 127)) // 64 * 64


def absprod(A, B):
    return oracle.dgemm(np.abs(A), np.abs(B))


def split_rep(opA, opB):
    """Representation error of the default (two-plane) split tier: each row-scaled
    operand is h + 2^-11 m, off by at most 2^-22 of its row's (column's) largest
    magnitude per element, so |dC| gains 2^-22 (max_row|A| * sum|B| + sum|A| * max_col|B|)
    (with a factor 2 margin)."""
    ra = np.abs(opA).max(axis=1, initial=0.0)[:, None]
    cb = np.abs(opB).max(axis=0, initial=0.0)[None, :]
    return 2.0 ** -21 * (ra * np.abs(opB).sum(axis=0)[None, :] +
                         np.abs(opA).sum(axis=1)[:, None] * cb)


def assert_bound(C, R, S, kfac, u, extra=3.0, lower=False, what="", extra_abs=0.0):
    bound = (gamma(max(kfac, 1), u) + extra * u) * S + extra_abs
    err = np.abs(C.astype(np.float64) - R)
    if lower:
        err = np.tril(err)
        bound = np.tril(bound) + np.triu(np.ones_like(bound), 1)
    bad = err > bound + 1e-300
    assert not bad.any(), (f"{what}: {bad.sum()} entries over the bound; worst ratio "
                           f"{np.max(err / np.maximum(bound, 1e-300)):.3g}")


def tier_dtype(tier):
    return np.float64 if tier in ("ozaki", "fp64") else np.float32


# ------------------------------------------------------------------- GEMM
@pytest.mark.parametrize("tier", ["plain", "split", "fp32", "fp64"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 65, 19), (128, 128, 64), (200, 131, 300)])
def test_gemm(h, tier, ta, tb, m, n, k):
    g = rng(m * 7 + n * 3 + k + 11 * ta + 13 * tb)
    dt = tier_dtype(tier)
    A = uniform(g, (k, m) if ta else (m, k), dtype=dt)
    B = uniform(g, (n, k) if tb else (k, n), dtype=dt)
    C0 = uniform(g, (m, n), dtype=dt)
    alpha, beta = 1.5, -0.5
    dA, lda = to_dev(A, A.shape[0] + 3)
    dB, ldb = to_dev(B, B.shape[0] + 1)
    dC, ldc = to_dev(C0, m + 2)
    h.gemm(tier, ta, tb, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta,
           dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, m, n, ldc)
    opA = A.T if ta else A
    opB = B.T if tb else B
    if tier == "plain":
        a16, b16 = oracle.round_f16(opA), oracle.round_f16(opB)
        R = alpha * oracle.dgemm(a16, b16) + beta * C0
        S = abs(alpha) * absprod(a16, b16) + abs(beta) * np.abs(C0)
        assert_bound(C, R, S, k, U32, what="plain gemm")
    else:
        R = alpha * oracle.dgemm(opA, opB) + beta * C0.astype(np.float64)
        S = abs(alpha) * absprod(opA, opB) + abs(beta) * np.abs(C0)
        u = U64 if tier == "fp64" else U32
        rep = abs(alpha) * split_rep(opA, opB) if tier == "split" else 0.0
        assert_bound(C, R, S, k, u, extra_abs=rep, what=f"{tier} gemm")


@pytest.mark.parametrize("m,n,k", [(2, 2, 1), (130, 258, 16), (256, 128, 17), (130, 66, 1000),
                                   (384, 256, 48)])
def test_gemm_fp64_nt_tight_ld(h, m, n, k):
    """FP64 NT products with leading dimensions equal to the row counts (even, 16-byte
    aligned columns): the operands end exactly where the matrices do."""
    g = rng(m + 3 * n + 7 * k)
    A, B = uniform(g, (m, k), dtype=np.float64), uniform(g, (n, k), dtype=np.float64)
    C0 = uniform(g, (m, n), dtype=np.float64)
    dA, lda = to_dev(A, m)
    dB, ldb = to_dev(B, n)
    dC, ldc = to_dev(C0, m)
    h.gemm("fp64", 0, 1, m, n, k, -1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.25,
           dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, m, n, ldc)
    R = -oracle.dgemm(A, B.T) + 0.25 * C0
    S = absprod(A, B.T) + 0.25 * np.abs(C0)
    assert_bound(C, R, S, k, U64, what="fp64 gemm nt")


# ------------------------------------------------------------------- SYRK
@pytest.mark.parametrize("tier", ["plain", "split", "fp32", "fp64"])
@pytest.mark.parametrize("n,k", [(1, 1), (7, 33), (129, 64), (300, 257), (513, 40)])
def test_syrk(h, tier, n, k):
    g = rng(n * 31 + k)
    dt = tier_dtype(tier)
    A = uniform(g, (n, k), dtype=dt)
    C0 = uniform(g, (n, n), dtype=dt)
    alpha, beta = (1.0, 0.0) if n % 2 else (-0.75, 1.25)
    dA, lda = to_dev(A, n + 5)
    dC, ldc = to_dev(C0, n + 1)
    h.syrk(tier, n, k, alpha, dA.data_ptr(), lda, beta, dC.data_ptr(), ldc, nb=256)
    h.synchronize()
    C = from_dev(dC, n, n, ldc)
    # strictly upper triangle untouched (BLAS semantics)
    assert np.array_equal(np.triu(C, 1), np.triu(C0, 1))
    if tier == "plain":
        a16 = oracle.round_f16(A)
        R = oracle.dsyrk(a16, alpha, beta, C0)
        S = abs(alpha) * absprod(a16, a16.T) + abs(beta) * np.abs(C0)
        assert_bound(C, R, S, k, U32, lower=True, what="plain syrk")
    else:
        R = oracle.dsyrk(A, alpha, beta, C0)
        S = abs(alpha) * absprod(A, A.T) + abs(beta) * np.abs(C0)
        u = U64 if tier == "fp64" else U32
        rep = abs(alpha) * split_rep(A, A.T) if tier == "split" else 0.0
        assert_bound(C, R, S, k, u, lower=True, what=f"{tier} syrk", extra_abs=rep)


def test_syrk_benchmark_size_plain(h):
    """SYRK benchmark shape n = k = 2048, nb = 256, alpha 1, beta 0 (plain tier)."""
    g = rng(0xC0F1)
    n = k = 2048
    A = uniform(g, (n, k))
    dA, lda = to_dev(A)
    dC, ldc = to_dev(np.zeros((n, n), np.float32))
    h.syrk("plain", n, k, 1.0, dA.data_ptr(), lda, 0.0, dC.data_ptr(), ldc, nb=256)
    h.synchronize()
    C = np.tril(from_dev(dC, n, n, ldc)).astype(np.float64)
    a16 = oracle.round_f16(A)
    R = np.tril(a16 @ a16.T)   # FP64 BLAS, pinned to the oracle by the tests of oracle.dsyrk
    S = np.tril(np.abs(a16) @ np.abs(a16).T)
    assert_bound(C, R, S, k, U32, lower=True, what="plain syrk 2048")
    e = oracle.errors(C, np.tril(A.astype(np.float64) @ A.T.astype(np.float64)), lower=True)
    assert e["normwise"] < 2.0 ** -10


# ------------------------------------------------------------------ SYR2K
@pytest.mark.parametrize("tier", ["plain", "split", "fp32", "fp64"])
@pytest.mark.parametrize("n,k,same_ld", [(5, 3, True), (130, 70, True), (257, 129, False)])
def test_syr2k(h, tier, n, k, same_ld):
    g = rng(n + 5 * k)
    dt = tier_dtype(tier)
    A, B = uniform(g, (n, k), dtype=dt), uniform(g, (n, k), dtype=dt)
    C0 = uniform(g, (n, n), dtype=dt)
    alpha, beta = 0.5, 2.0
    dA, lda = to_dev(A, n + 2)
    dB, ldb = to_dev(B, n + 2 if same_ld else n + 7)
    dC, ldc = to_dev(C0, n)
    h.syr2k(tier, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, n, n, ldc)
    if tier == "plain":
        A, B = oracle.round_f16(A), oracle.round_f16(B)
    R = oracle.dsyr2k(A, B, alpha, beta, C0)
    S = abs(alpha) * (absprod(A, B.T) + absprod(B, A.T)) + abs(beta) * np.abs(C0)
    u = U64 if tier == "fp64" else U32
    rep = abs(alpha) * (split_rep(A, B.T) + split_rep(B, A.T)) if tier == "split" else 0.0
    assert_bound(C, R, S, 2 * k, u, lower=True, what=f"{tier} syr2k", extra_abs=rep)


# ------------------------------------------------------------------- TRMM
@pytest.mark.parametrize("tier", ["plain", "split", "fp32", "fp64"])
@pytest.mark.parametrize("m,n", [(1, 3), (100, 37), (300, 129)])
def test_trmm(h, tier, m, n):
    g = rng(m * 3 + n)
    dt = tier_dtype(tier)
    L = uniform(g, (m, m), dtype=dt)       # strictly upper part is garbage: ignored
    B = uniform(g, (m, n), dtype=dt)
    dL, lda = to_dev(L)
    dB, ldb = to_dev(B, m + 3)
    dC, ldc = to_dev(np.full((m, n), 7.0, dt), m + 1)
    h.trmm(tier, m, n, -2.0, dL.data_ptr(), lda, dB.data_ptr(), ldb, dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, m, n, ldc)
    Lt = np.tril(L)
    if tier == "plain":
        Lt, B = oracle.round_f16(Lt), oracle.round_f16(B)
    R = oracle.dtrmm(L if tier != "plain" else Lt, B, -2.0)
    S = 2.0 * absprod(np.tril(Lt), B)
    u = U64 if tier == "fp64" else U32
    rep = 2.0 * split_rep(np.tril(Lt), B) if tier == "split" else 0.0
    assert_bound(C, R, S, m, u, what=f"{tier} trmm", extra_abs=rep)


# ------------------------------------------------------------------- SYMM
@pytest.mark.parametrize("tier", ["plain", "split", "fp32", "fp64"])
@pytest.mark.parametrize("m,n", [(2, 2), (129, 40), (260, 300)])
def test_symm(h, tier, m, n):
    g = rng(m * 17 + n)
    dt = tier_dtype(tier)
    A = uniform(g, (m, m), dtype=dt)        # only the lower triangle is the operand
    B, C0 = uniform(g, (m, n), dtype=dt), uniform(g, (m, n), dtype=dt)
    dA, lda = to_dev(A, m + 4)
    A_before = dA.clone()
    dB, ldb = to_dev(B)
    dC, ldc = to_dev(C0)
    h.symm(tier, m, n, 1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.5, dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, m, n, ldc)
    assert torch.equal(A_before.nan_to_num(123.0), dA.nan_to_num(123.0))   # A never modified
    Asym = np.tril(A) + np.tril(A, -1).T
    if tier == "plain":
        Asym, B = oracle.round_f16(Asym), oracle.round_f16(B)
    R = oracle.dsymm(Asym, B, 1.0, 0.5, C0)
    S = absprod(Asym, B) + 0.5 * np.abs(C0)
    u = U64 if tier == "fp64" else U32
    rep = split_rep(Asym, B) if tier == "split" else 0.0
    assert_bound(C, R, S, m, u, what=f"{tier} symm", extra_abs=rep)


# ------------------------------------------------------------------- TRSM
TRSM_TOL = {"plain": 2.0 ** -9, "split": 1e-6, "fp32": 1e-6, "fp64": 1e-13, "ozaki": 1e-13}


@pytest.mark.parametrize("tier", ["plain", "split", "fp32", "fp64", "ozaki"])
@pytest.mark.parametrize("m,n,nb", [(1, 1, 256), (7, 300, 256), (64, 2048, 256), (130, 2048, 512),
                                    (33, 4100, 256), (8200, 520, 256)])
def test_trsm(h, tier, m, n, nb):
    g = rng(m * 5 + n + nb)
    dt = tier_dtype(tier)
    L = lower_tri_dominant(g, n, dt)
    L[np.triu_indices(n, 1)] = np.nan          # the strict upper triangle is never read
    B = uniform(g, (m, n), dtype=dt)
    dL, lda = to_dev(L)
    dB, ldb = to_dev(B, m + 1)
    h.trsm(tier, m, n, dL.data_ptr(), lda, dB.data_ptr(), ldb, nb=nb)
    h.synchronize()
    X = from_dev(dB, m, n, ldb)
    Lc = np.tril(np.nan_to_num(L))
    R = oracle.dtrsm_rltn(Lc, B)
    e = oracle.errors(X, R)
    assert e["normwise"] < TRSM_TOL[tier] and e["maxrel"] < 4 * TRSM_TOL[tier], (tier, e)
    if tier == "plain" and n % 2048 == 0:
        # the reference's own recursion with FP16-rounded updates, emulated exactly in FP64
        em = oracle.tc_trsm_plain_emul(Lc, B, nb)
        assert oracle.errors(X, em)["normwise"] < 1e-5


def test_trsm_split_at_least_as_accurate_as_fp32(h):
    g = rng(0xC0F2)
    m, n = 64, 4096
    L = lower_tri_dominant(g, n)
    B = uniform(g, (m, n))
    R = oracle.dtrsm_rltn(L, B)
    errs = {}
    for tier in ("split", "fp32", "plain"):
        dL, lda = to_dev(L)
        dB, ldb = to_dev(B)
        h.trsm(tier, m, n, dL.data_ptr(), lda, dB.data_ptr(), ldb, nb=256)
        h.synchronize()
        errs[tier] = oracle.errors(from_dev(dB, m, n, ldb), R)
    assert errs["split"]["normwise"] <= 1.5 * errs["fp32"]["normwise"], errs
    assert errs["split"]["maxrel"] <= 1.5 * errs["fp32"]["maxrel"], errs
    assert errs["plain"]["normwise"] > errs["split"]["normwise"], errs


@pytest.mark.parametrize("tier", ["split", "plain"])
def test_trsm_refined(h, tier):
    g = rng(77)
    m, n = 40, 2048
    L = lower_tri_dominant(g, n)
    B = uniform(g, (m, n))
    R = oracle.dtrsm_rltn(L, B)
    dL, lda = to_dev(L)
    dB, ldb = to_dev(B)
    h.trsm(tier, m, n, dL.data_ptr(), lda, dB.data_ptr(), ldb)
    e0 = oracle.errors(from_dev(dB, m, n, ldb), R)["normwise"]
    dB2, _ = to_dev(B)
    h.trsm_refined(tier, m, n, dL.data_ptr(), lda, dB2.data_ptr(), ldb)
    h.synchronize()
    e1 = oracle.errors(from_dev(dB2, m, n, ldb), R)["normwise"]
    assert e1 <= e0 and e1 < 1e-6, (e0, e1)


# ------------------------------------------------------------------ Ozaki
def _oz_slices(X, S):
    return oracle.ozaki_slice(X, S)


@pytest.mark.parametrize("S,extra", [(3, 0), (6, 0), (7, 0), (9, 0), (6, 1), (8, 1), (9, 1)])
@pytest.mark.parametrize("m,n,k,phi", [(17, 9, 33, 0.5), (70, 130, 200, 2.0), (64, 64, 64, 0.5)])
def test_ozaki_gemm_bit_exact(h, S, extra, m, n, k, phi):
    g = rng(S * 101 + m + n + k)
    A, B = spread(g, (m, k), phi), spread(g, (k, n), phi)
    C0 = spread(g, (m, n), phi)
    h.set_ozaki_slices(S)
    h.set_ozaki_extra(extra)
    dA, lda = to_dev(A, m + 1)
    dB, ldb = to_dev(B, k)
    dC, ldc = to_dev(C0, m)
    h.gemm("ozaki", 0, 0, m, n, k, 1.25, dA.data_ptr(), lda, dB.data_ptr(), ldb, -0.5,
           dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, m, n, ldc)
    sa, ea = _oz_slices(A, S)
    sb, eb = _oz_slices(B.T.copy(), S)
    R = oracle.ozaki_gemm(sa, ea, sb, eb, 1.25, -0.5, C0, kchunk=_ozaki_kchunk(S), ngroups=S + extra)
    h.set_ozaki_extra(0)
    h.set_ozaki_slices(7)
    assert np.array_equal(C, R), np.max(np.abs(C - R))


@pytest.mark.parametrize("op", ["gemm", "syrk"])
def test_ozaki_k_zero_scales_c(h, op):
    """k = 0: C := beta C exactly (the lower triangle for SYRK); the operand buffers hold
    only NaN, so any value read from them would show in C."""
    g = rng(4242)
    n = 70
    C0 = spread(g, (n, n), 0.5)
    dA, lda = to_dev(np.zeros((n, 0)), n)
    dC, ldc = to_dev(C0, n + 1)
    if op == "gemm":
        dB, ldb = to_dev(np.zeros((0, n)), 1)
        h.gemm("ozaki", 0, 0, n, n, 0, 1.25, dA.data_ptr(), lda, dB.data_ptr(), ldb, -0.5,
               dC.data_ptr(), ldc)
    else:
        h.syrk("ozaki", n, 0, 1.25, dA.data_ptr(), lda, -0.5, dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, n, n, ldc)
    R = -0.5 * C0
    if op == "syrk":
        R = np.tril(R) + np.triu(C0, 1)
    assert np.array_equal(C, R)


@pytest.mark.parametrize("S,extra", [(9, 0), (8, 1)])
def test_ozaki_long_k_bit_exact(h, S, extra):
    """K past one exactness chunk with beta = 0 runs as one launch per chunk (each chunk's
    FP64 sum kept in C between them): the chunked kernel's bits.  GEMM with three chunks,
    SYR2K with two chunks of each product."""
    kc = _ozaki_kchunk(S)
    g = rng(9000 + S)
    h.set_ozaki_slices(S)
    h.set_ozaki_extra(extra)
    try:
        m, n, k = 70, 33, 2 * kc + 100
        A, B = spread(g, (m, k), 0.5), spread(g, (k, n), 0.5)
        dA, lda = to_dev(A)
        dB, ldb = to_dev(B)
        dC, ldc = to_dev(np.full((m, n), np.nan))
        h.gemm("ozaki", 0, 0, m, n, k, 1.5, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.0,
               dC.data_ptr(), ldc)
        h.synchronize()
        sa, ea = _oz_slices(A, S)
        sb, eb = _oz_slices(B.T.copy(), S)
        R = oracle.ozaki_gemm(sa, ea, sb, eb, 1.5, 0.0, None, kchunk=_ozaki_kchunk(S), ngroups=S + extra)
        C = from_dev(dC, m, n, ldc)
        assert np.array_equal(C, R), np.max(np.abs(C - R))
        # SYR2K: each product chunked at kchunk; the oracle's concatenated K pads each half
        # to a multiple of kchunk so that the chunk boundaries fall in the same places
        nn, k2 = 40, kc + 64
        A2, B2 = spread(g, (nn, k2), 2.0), spread(g, (nn, k2), 2.0)
        dA2, lda2 = to_dev(A2)
        dB2, ldb2 = to_dev(B2)
        dC2, ldc2 = to_dev(np.full((nn, nn), np.nan))
        h.syr2k("ozaki", nn, k2, 1.0, dA2.data_ptr(), lda2, dB2.data_ptr(), ldb2, 0.0, dC2.data_ptr(), ldc2)
        h.synchronize()
        kp = 2 * _ozaki_kchunk(S)
        Ap = np.zeros((nn, 2 * kp)); Ap[:, :k2] = A2; Ap[:, kp:kp + k2] = B2
        Bp = np.zeros((nn, 2 * kp)); Bp[:, :k2] = B2; Bp[:, kp:kp + k2] = A2
        s1, e1 = _oz_slices(Ap, S)
        s2, e2 = _oz_slices(Bp, S)
        R2 = oracle.ozaki_gemm(s1, e1, s2, e2, 1.0, 0.0, None, kchunk=_ozaki_kchunk(S), ngroups=S + extra)
        C2 = from_dev(dC2, nn, nn, ldc2)
        assert np.array_equal(np.tril(C2), np.tril(R2)), np.max(np.abs(np.tril(C2) - np.tril(R2)))
    finally:
        h.set_ozaki_extra(0)
        h.set_ozaki_slices(7)


def test_ozaki_extra_group_arguments(h, tbm):
    h.set_ozaki_slices(4)
    h.set_ozaki_extra(1)                      # accepted; the call then refuses nslices 4
    z = torch.zeros(16, device="cuda", dtype=torch.float64)
    with pytest.raises(tbm.TensorBLASError):
        h.gemm("ozaki", 0, 0, 4, 4, 4, 1.0, z.data_ptr(), 4, z.data_ptr(), 4, 0.0, z.data_ptr(), 4)
    h.set_ozaki_extra(0)
    h.set_ozaki_slices(7)
    with pytest.raises(tbm.TensorBLASError):
        h.set_ozaki_extra(2)


def test_ozaki_routines_bit_exact(h):
    g = rng(4242)
    S = 7
    h.set_ozaki_slices(S)
    n, k = 70, 45
    A, B = spread(g, (n, k), 2.0), spread(g, (n, k), 2.0)
    C0 = spread(g, (n, n), 0.5)
    # SYRK
    dA, lda = to_dev(A)
    dC, ldc = to_dev(C0)
    h.syrk("ozaki", n, k, 1.0, dA.data_ptr(), lda, 0.5, dC.data_ptr(), ldc)
    h.synchronize()
    sa, ea = _oz_slices(A, S)
    R = oracle.ozaki_gemm(sa, ea, sa, ea, 1.0, 0.5, C0, kchunk=_ozaki_kchunk(S))
    C = from_dev(dC, n, n, ldc)
    assert np.array_equal(np.tril(C), np.tril(R))
    assert np.array_equal(np.triu(C, 1), np.triu(C0, 1))
    # SYR2K: the two products run as one concatenated K (each half zero-padded to 64)
    dB, ldb = to_dev(B)
    dC, ldc = to_dev(C0)
    h.syr2k("ozaki", n, k, 1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.0, dC.data_ptr(), ldc)
    h.synchronize()
    kp = 64
    Ap = np.zeros((n, 2 * kp)); Ap[:, :k] = A; Ap[:, kp:kp + k] = B
    Bp = np.zeros((n, 2 * kp)); Bp[:, :k] = B; Bp[:, kp:kp + k] = A
    s1, e1 = _oz_slices(Ap, S)
    s2, e2 = _oz_slices(Bp, S)
    R = oracle.ozaki_gemm(s1, e1, s2, e2, 1.0, 0.0, None, kchunk=kp)
    assert np.array_equal(np.tril(from_dev(dC, n, n, ldc)), np.tril(R))
    # TRMM (upper part of L ignored) and SYMM (A stored lower)
    m, nn = 50, 31
    L = spread(g, (m, m), 0.5)
    Bm = spread(g, (m, nn), 0.5)
    dL, ldl = to_dev(L)
    dB2, ldb2 = to_dev(Bm)
    dC2, ldc2 = to_dev(np.zeros((m, nn)))
    h.trmm("ozaki", m, nn, 2.0, dL.data_ptr(), ldl, dB2.data_ptr(), ldb2, dC2.data_ptr(), ldc2)
    h.synchronize()
    sl, el = _oz_slices(np.tril(L), S)
    sbb, ebb = _oz_slices(Bm.T.copy(), S)
    R = oracle.ozaki_gemm(sl, el, sbb, ebb, 2.0, 0.0, None, kchunk=_ozaki_kchunk(S))
    assert np.array_equal(from_dev(dC2, m, nn, ldc2), R)
    dC3, _ = to_dev(np.ones((m, nn)))
    h.symm("ozaki", m, nn, 1.0, dL.data_ptr(), ldl, dB2.data_ptr(), ldb2, 3.0, dC3.data_ptr(), ldc2)
    h.synchronize()
    Asym = np.tril(L) + np.tril(L, -1).T
    ss, es = _oz_slices(Asym, S)
    R = oracle.ozaki_gemm(ss, es, sbb, ebb, 1.0, 3.0, np.ones((m, nn)), kchunk=_ozaki_kchunk(S))
    assert np.array_equal(from_dev(dC3, m, nn, ldc2), R)


@pytest.mark.parametrize("S,extra,n,k,beta", [
    (8, 1, 300, 130, 0.0),       # one exactness chunk per product: two launches, sum kept in C
    (8, 1, 300, 130, 0.5),       # beta != 0: both products in one launch of the chunked kernel
    (9, 1, 257, 64, 0.0),
    (10, 0, 129, 700, 0.0),
])
def test_ozaki_syr2k_paths_bit_exact(h, S, extra, n, k, beta):
    g = rng(S * 7 + n + k)
    A, B = spread(g, (n, k), 2.0), spread(g, (n, k), 2.0)
    C0 = spread(g, (n, n), 0.5)
    h.set_ozaki_slices(S)
    h.set_ozaki_extra(extra)
    dA, lda = to_dev(A, n + 3)
    dB, ldb = to_dev(B, n + 1)
    dC, ldc = to_dev(C0, n + 2)
    h.syr2k("ozaki", n, k, 0.75, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, n, n, ldc)
    h.set_ozaki_extra(0)
    h.set_ozaki_slices(7)
    # the two products as one concatenated K, each half zero-padded to a multiple of 64
    kp = (k + 63) // 64 * 64
    Ap = np.zeros((n, 2 * kp)); Ap[:, :k] = A; Ap[:, kp:kp + k] = B
    Bp = np.zeros((n, 2 * kp)); Bp[:, :k] = B; Bp[:, kp:kp + k] = A
    s1, e1 = _oz_slices(Ap, S)
    s2, e2 = _oz_slices(Bp, S)
    # each product is one exactness chunk (K within kchunk)
    assert kp <= _ozaki_kchunk(S)
    R = oracle.ozaki_gemm(s1, e1, s2, e2, 0.75, beta, C0, kchunk=kp, ngroups=S + extra)
    assert np.array_equal(np.tril(C), np.tril(R)), np.max(np.abs(np.tril(C - R)))
    assert np.array_equal(np.triu(C, 1), np.triu(C0, 1))


def test_ozaki_syrk_past_one_chunk_bit_exact(h):
    g = rng(515)
    S, n, k = 10, 96, 13400
    A = spread(g, (n, k), 0.5)
    C0 = spread(g, (n, n), 0.5)
    h.set_ozaki_slices(S)
    dA, lda = to_dev(A)
    dC, ldc = to_dev(C0)
    h.syrk("ozaki", n, k, 1.0, dA.data_ptr(), lda, -1.0, dC.data_ptr(), ldc)
    h.synchronize()
    C = from_dev(dC, n, n, ldc)
    h.set_ozaki_slices(7)
    sa, ea = _oz_slices(A, S)
    R = oracle.ozaki_gemm(sa, ea, sa, ea, 1.0, -1.0, C0, kchunk=_ozaki_kchunk(S))
    assert np.array_equal(np.tril(C), np.tril(R)), np.max(np.abs(np.tril(C - R)))


def test_ozaki_accuracy_vs_fp64(h):
    g = rng(0xC0F4)
    n, k = 256, 512
    for phi in (0.5, 2.0):
        A = spread(g, (n, k), phi)
        B = spread(g, (n, k), phi)
        R = oracle.dsyr2k(A, B)
        for S, tol in ((6, 1e-10), (8, 1e-14)):
            h.set_ozaki_slices(S)
            dA, lda = to_dev(A)
            dB, ldb = to_dev(B)
            dC, ldc = to_dev(np.zeros((n, n)))
            h.syr2k("ozaki", n, k, 1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.0,
                    dC.data_ptr(), ldc)
            h.synchronize()
            e = oracle.errors(np.tril(from_dev(dC, n, n, ldc)), np.tril(R), lower=True)
            assert e["maxrel"] < tol, (phi, S, e)
    h.set_ozaki_slices(7)


# ---------------------------------------------------- conversions and wrappers
def test_s2h_bit_exact(h):
    g = rng(3)
    m, n = 300, 77
    X = (g.standard_normal((m, n)) * np.exp2(g.integers(-30, 20, (m, n)))).astype(np.float32)
    X[0, 0], X[1, 0], X[2, 0] = 70000.0, -0.0, 2.0 ** -25
    dX, ldx = to_dev(X, m + 3)
    dH = torch.zeros(n * m, dtype=torch.float16, device="cuda")
    h.s2h(m, n, dX.data_ptr(), ldx, dH.data_ptr(), m)
    h.synchronize()
    got = dH.cpu().numpy().view(np.uint16).reshape(n, m).T
    assert np.array_equal(got, oracle.f32_to_f16_bits(X))
    # row-major source (util/util.cu:137-145)
    dR = torch.from_numpy(np.ascontiguousarray(X)).cuda()
    h.s2h_transpose(m, n, dR.data_ptr(), dH.data_ptr())
    h.synchronize()
    got = dH.cpu().numpy().view(np.uint16).reshape(n, m).T
    assert np.array_equal(got, oracle.f32_to_f16_bits(X))


@pytest.mark.parametrize("f64", [False, True])
def test_copy_lower_to_upper(h, f64):
    g = rng(5)
    n = 150
    A = uniform(g, (n, n), dtype=np.float64 if f64 else np.float32)
    dA, lda = to_dev(A, n + 2)
    h.copy_lower_to_upper(f64, n, dA.data_ptr(), lda)
    h.synchronize()
    got = from_dev(dA, n, n, lda)
    assert np.array_equal(got, np.tril(A) + np.tril(A, -1).T)


def test_reference_python_binding():
    import TensorBLAS
    g = rng(9)
    A = torch.from_numpy(uniform(g, (300, 120))).cuda()
    C = TensorBLAS.tc_syrk(A).cpu().numpy().astype(np.float64)
    a16 = oracle.round_f16(A.cpu().numpy())
    R = a16 @ a16.T
    assert np.allclose(C, C.T)
    assert_bound(C, R, np.abs(a16) @ np.abs(a16).T, 120, U32, what="tc_syrk_wrapper")
    Ct = TensorBLAS.tc_syrk(A, transpose=True).cpu().numpy()
    Rt = a16.T @ a16
    assert oracle.errors(Ct, Rt)["normwise"] < 1e-6
    n, m = 260, 90
    L = torch.from_numpy(lower_tri_dominant(g, n)).cuda()
    Bt = torch.from_numpy(uniform(g, (m, n))).cuda()
    X = TensorBLAS.tc_trsm(L, Bt).cpu().numpy()
    R = oracle.dtrsm_rltn(L.cpu().numpy(), Bt.cpu().numpy())
    assert oracle.errors(X, R)["normwise"] < 1e-6


def test_error_codes(h, tbm):
    with pytest.raises(tbm.TensorBLASError) as ei:
        h.syrk("split", 10, 10, 1.0, 0, 10, 0.0, 0, 10)       # null pointers
    assert ei.value.status == 1
    with pytest.raises(tbm.TensorBLASError):
        h.syrk("split", 10, 10, 1.0, 1, 5, 0.0, 1, 10)        # lda < n
    with pytest.raises(tbm.TensorBLASError):
        h.trsm("plain", 10, 10, 1, 10, 1, 10, nb=0)           # nb <= 0
    with pytest.raises(tbm.TensorBLASError):
        h.trsm_refined("fp64", 10, 10, 1, 10, 1, 10)          # FP32 tiers only
    with pytest.raises(tbm.TensorBLASError):
        h.set_ozaki_slices(11)
    h.syrk("plain", 0, 10, 1.0, 0, 1, 0.0, 0, 1)              # empty: no-op
    h.synchronize()


# --------------------------------------------- split tier with three FP16 planes
# h + 2^-11 m + 2^-22 l is exact for FP32 inputs, so the three-plane tier keeps a
# componentwise bound: FP32 accumulation plus 2^-20 = 16u for the products of weight
# < 2^-22 it does not form.
@pytest.fixture(scope="module")
def h3(tbm):
    hd = tbm.Handle()
    hd.set_split_planes(3)
    yield hd
    hd.close()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (37, 65, 19), (200, 131, 300), (256, 256, 1024)])
def test_gemm_split3(h3, ta, tb, m, n, k):
    g = rng(m * 3 + n * 5 + k + 7 * ta + 11 * tb)
    A = uniform(g, (k, m) if ta else (m, k))
    B = uniform(g, (n, k) if tb else (k, n))
    C0 = uniform(g, (m, n))
    dA, lda = to_dev(A, A.shape[0] + 1)
    dB, ldb = to_dev(B, B.shape[0] + 2)
    dC, ldc = to_dev(C0, m)
    h3.gemm("split", ta, tb, m, n, k, 1.5, dA.data_ptr(), lda, dB.data_ptr(), ldb, -0.5,
            dC.data_ptr(), ldc)
    h3.synchronize()
    C = from_dev(dC, m, n, ldc)
    opA, opB = (A.T if ta else A), (B.T if tb else B)
    R = 1.5 * oracle.dgemm(opA, opB) - 0.5 * C0.astype(np.float64)
    S = 1.5 * absprod(opA, opB) + 0.5 * np.abs(C0)
    assert_bound(C, R, S, k, U32, extra=3.0 + 16.0, what="split3 gemm")


@pytest.mark.parametrize("n,k", [(129, 64), (300, 257)])
def test_syrk_split3(h3, n, k):
    g = rng(n + 17 * k)
    A = uniform(g, (n, k))
    dA, lda = to_dev(A, n)
    dC, ldc = to_dev(np.zeros((n, n), np.float32), n)
    h3.syrk("split", n, k, 1.0, dA.data_ptr(), lda, 0.0, dC.data_ptr(), ldc, nb=256)
    h3.synchronize()
    C = from_dev(dC, n, n, ldc)
    assert_bound(C, oracle.dsyrk(A), absprod(A, A.T), k, U32, extra=3.0 + 16.0, lower=True,
                 what="split3 syrk")


@pytest.mark.parametrize("m,n,nb", [(7, 300, 256), (64, 2048, 256), (33, 4100, 256)])
def test_trsm_split3(h3, m, n, nb):
    g = rng(m * 5 + n + nb)
    L = lower_tri_dominant(g, n)
    L[np.triu_indices(n, 1)] = np.nan
    B = uniform(g, (m, n))
    dL, lda = to_dev(L)
    dB, ldb = to_dev(B, m + 1)
    h3.trsm("split", m, n, dL.data_ptr(), lda, dB.data_ptr(), ldb, nb=nb)
    h3.synchronize()
    X = from_dev(dB, m, n, ldb)
    e = oracle.errors(X, oracle.dtrsm_rltn(np.tril(np.nan_to_num(L)), B))
    assert e["normwise"] < TRSM_TOL["split"] and e["maxrel"] < 4 * TRSM_TOL["split"], e


def test_trsm_split3_vs_fp32(h3):
    g = rng(0xC0F2)
    m, n = 64, 4096
    L = lower_tri_dominant(g, n)
    B = uniform(g, (m, n))
    R = oracle.dtrsm_rltn(L, B)
    errs = {}
    for tier in ("split", "fp32"):
        dL, lda = to_dev(L)
        dB, ldb = to_dev(B)
        h3.trsm(tier, m, n, dL.data_ptr(), lda, dB.data_ptr(), ldb, nb=256)
        h3.synchronize()
        errs[tier] = oracle.errors(from_dev(dB, m, n, ldb), R)
    assert errs["split"]["normwise"] <= 1.5 * errs["fp32"]["normwise"], errs
    assert errs["split"]["maxrel"] <= 1.5 * errs["fp32"]["maxrel"], errs
