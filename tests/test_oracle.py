"""Pins the CPU oracle before anything is checked against it.

The reference ships no golden vectors or thresholds (its drivers only print
errors, e.g. testing/testing_syrk.cu:152), so the oracle is pinned by:
  * known answers of the chunk planners (util/mat_size.cu:5-55),
  * an independent library: OpenBLAS through scipy.linalg.blas for every FP64
    BLAS-3 routine, and numpy's IEEE float16 cast for the FP16 conversion,
  * exact special cases (FP16-exact recursions, identity triangles),
  * exact integer / rational arithmetic for the Ozaki slicing.
"""
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg.blas as sb

import oracle


def rng(seed):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------- planners
@pytest.mark.parametrize("n,expect", [
    (2048, [2048]), (16384, [16384]), (32768, [16384, 16384]), (20000, [16384, 3616]),
    (8192, [8192]), (24576, [16384, 8192]), (40960, [16384, 16384, 8192]), (1, [1]),
    (0, []), (8191, [8191]), (16385, [16384, 1]),
])
def test_find_mat_size_syrk(n, expect):
    assert oracle.find_mat_size_syrk(n) == expect


@pytest.mark.parametrize("n,expect", [
    (16384, [16384]), (20000, [16384, 2048, 1568]), (2048, [2048]), (6144, [4096, 2048]),
    (14336, [8192, 4096, 2048]), (300, [300]), (30720, [16384, 8192, 4096, 2048]),
])
def test_find_mat_size_trsm(n, expect):
    assert oracle.find_mat_size_trsm(n) == expect


def test_planner_domain():
    # the reference's power tables sum to 188416 (SYRK) and 161792 (TRSM); a remainder
    # at or above the floor after the last power reads powers[-1]
    # (util/mat_size.cu:12,20).  The restatement reports -1 there.
    out = np.zeros(16, dtype=np.int64)
    L = oracle.lib()
    assert L.orc_find_mat_size_syrk(188416 + 8191, out.ctypes.data) == 13
    assert L.orc_find_mat_size_syrk(188416 + 8192, out.ctypes.data) == -1
    assert L.orc_find_mat_size_trsm(161792 + 2047, out.ctypes.data) == 13
    assert L.orc_find_mat_size_trsm(161792 + 2048, out.ctypes.data) == -1
    assert L.orc_find_mat_size_trsm(161792, out.ctypes.data) == 12


def test_planner_invariants():
    for n in list(range(0, 70000, 997)) + [188416, 161792]:
        for f, floor in ((oracle.find_mat_size_syrk, 8192), (oracle.find_mat_size_trsm, 2048)):
            if f is oracle.find_mat_size_trsm and n > 161792:
                continue
            ch = f(n)
            assert sum(ch) == n
            assert all(c >= floor for c in ch[:-1])
            assert ch == sorted(ch, reverse=True)


# ------------------------------------------------------------- conversions
def _check_f16_block(bits):
    x = bits.view(np.float32)
    mine = oracle.f32_to_f16_bits(x)
    ref = x.astype(np.float16).view(np.uint16)
    nan = np.isnan(x)
    assert np.array_equal(mine[~nan], ref[~nan])
    # NaN stays NaN (payload is not part of the contract)
    m = mine[nan]
    assert np.all((m & 0x7C00) == 0x7C00) and np.all((m & 0x3FF) != 0)


def test_f16_rne_special_values():
    vals = np.array([0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, -65520.0, 1e9,
                     np.inf, -np.inf, 2.0 ** -14, 2.0 ** -24, 2.0 ** -25, 2.0 ** -25 * 1.0001,
                     2.0 ** -26, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 6.1e-5, 3.0e-8],
                    dtype=np.float32)
    _check_f16_block(vals.view(np.uint32))
    assert oracle.lib().orc_f32_to_f16(65520.0) == 0x7C00
    assert oracle.lib().orc_f32_to_f16(-0.0) == 0x8000
    assert oracle.lib().orc_f32_to_f16(1.0 + 2.0 ** -11) == 0x3C00      # tie -> even
    assert oracle.lib().orc_f32_to_f16(1.0 + 3 * 2.0 ** -11) == 0x3C02  # tie -> even (up)


def test_f16_rne_all_exponents_sampled():
    # every exponent/sign, 8192 mantissas each (including the tie and boundary patterns)
    g = rng(1)
    man = np.concatenate([g.integers(0, 1 << 23, 8000, dtype=np.uint32),
                          np.array([0, 1, 0x1000, 0x0FFF, 0x1001, 0x2000, 0x3000, 0x7FFFFF,
                                    0x7FE000, 0x7FF000, 0x400000], dtype=np.uint32)])
    ex = np.arange(256, dtype=np.uint32)
    bits = (ex[:, None] << 23 | man[None, :]).ravel()
    bits = np.concatenate([bits, bits | 0x80000000])
    _check_f16_block(bits)


def test_f16_rne_rounding_bits_exhaustive():
    # RNE depends on the exponent, the 13 dropped mantissa bits and the kept LSB (and
    # carries through the kept bits): every exponent, sign and dropped-bit pattern,
    # with 16 kept-mantissa patterns that include the carry boundaries.
    low = np.arange(1 << 13, dtype=np.uint32)
    kept = np.array([0, 1, 2, 3, 0x155, 0x2AA, 0x1FF, 0x200, 0x3FE, 0x3FF, 0x100, 0x0FF,
                     0x333, 0x0F0, 0x3C3, 0x001 | 0x200], dtype=np.uint32) << 13
    mant = (kept[:, None] | low[None, :]).ravel()
    for e in range(256):
        bits = (np.uint32(e) << np.uint32(23)) | mant
        _check_f16_block(np.concatenate([bits, bits | np.uint32(0x80000000)]))


@pytest.mark.skipif(os.environ.get("TB_EXHAUSTIVE") != "1",
                    reason="all 2^32 patterns take ~7 min; set TB_EXHAUSTIVE=1")
def test_f16_rne_exhaustive():
    step = 1 << 26
    for start in range(0, 1 << 32, step):
        _check_f16_block(np.arange(start, start + step, dtype=np.uint64).astype(np.uint32))


def test_f16_roundtrip_exact_halves():
    h = np.arange(0, 1 << 16, dtype=np.uint16)
    f = h.view(np.float16).astype(np.float32)
    ok = ~np.isnan(f)
    assert np.array_equal(oracle.f32_to_f16_bits(f[ok]), h[ok])
    L = oracle.lib()
    for v in h[ok][::97]:
        assert L.orc_f16_to_f32(int(v)) == float(np.uint16(v).view(np.float16))


def test_bf16_rne():
    g = rng(2)
    x = g.standard_normal(20000).astype(np.float32) * np.exp2(g.integers(-60, 60, 20000)).astype(np.float32)
    L = oracle.lib()
    for v in x[:3000]:
        b = L.orc_f32_to_bf16(float(v))
        # nearest bf16 by brute force over the two neighbours
        u = np.float32(v).view(np.uint32)
        lo = (int(u) >> 16) << 16
        hi = lo + (1 << 16)
        cand = [np.uint32(lo).view(np.float32), np.uint32(hi & 0xFFFFFFFF).view(np.float32)]
        d = [abs(float(c) - float(v)) for c in cand]
        if d[0] < d[1]:
            want = lo >> 16
        elif d[1] < d[0]:
            want = (hi >> 16) & 0xFFFF
        else:
            want = (lo >> 16) if ((lo >> 16) & 1) == 0 else ((hi >> 16) & 0xFFFF)
        assert b == want


# ------------------------------------------------------------------ BLAS-3
@pytest.mark.parametrize("n,k", [(1, 1), (3, 5), (7, 64), (64, 3), (65, 33)])
def test_dsyrk_vs_openblas(n, k):
    g = rng(n * 100 + k)
    A = g.uniform(-1, 1, (n, k))
    C0 = g.uniform(-1, 1, (n, n))
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        mine = oracle.dsyrk(A, alpha, beta, C0)
        ref = sb.dsyrk(alpha, A, beta=beta, c=np.asfortranarray(C0), lower=1)
        assert np.allclose(np.tril(mine), np.tril(ref), rtol=1e-13, atol=1e-13)
        assert np.array_equal(np.triu(mine, 1), np.triu(C0, 1))


@pytest.mark.parametrize("n,k", [(1, 1), (5, 3), (33, 17), (64, 64)])
def test_dsyr2k_vs_openblas(n, k):
    g = rng(n + 7 * k)
    A, B = g.uniform(-1, 1, (n, k)), g.standard_normal((n, k))
    C0 = g.uniform(-1, 1, (n, n))
    mine = oracle.dsyr2k(A, B, 0.75, 0.5, C0)
    ref = sb.dsyr2k(0.75, A, B, beta=0.5, c=np.asfortranarray(C0), lower=1)
    assert np.allclose(np.tril(mine), np.tril(ref), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("m,n", [(1, 1), (4, 9), (31, 7), (64, 65)])
def test_dtrmm_vs_openblas(m, n):
    g = rng(m * 3 + n)
    L = g.uniform(-1, 1, (m, m))          # upper part is garbage: must be ignored
    B = g.uniform(-1, 1, (m, n))
    mine = oracle.dtrmm(L, B, 1.5)
    ref = sb.dtrmm(1.5, np.asfortranarray(L), np.asfortranarray(B), side=0, lower=1)
    assert np.allclose(mine, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("m,n", [(1, 1), (3, 3), (64, 17), (300, 64)])
def test_dtrsm_vs_openblas(m, n):
    g = rng(m + 11 * n)
    L = g.uniform(-1, 1, (n, n))
    L[np.diag_indices(n)] = n + g.uniform(0, 1, n)
    B = g.uniform(-1, 1, (m, n))
    mine = oracle.dtrsm_rltn(L, B)
    ref = sb.dtrsm(1.0, np.asfortranarray(L), np.asfortranarray(B), side=1, lower=1, trans_a=1)
    assert np.allclose(mine, ref, rtol=1e-12, atol=1e-15)
    # left/lower/N form: L X = B^T  <=>  X^T L^T = B
    mine_l = oracle.dtrsm_llnn(L, B.T)
    assert np.allclose(mine_l, ref.T, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("m,n", [(1, 1), (6, 2), (40, 33)])
def test_dsymm_vs_openblas(m, n):
    g = rng(m * 13 + n)
    A = g.uniform(-1, 1, (m, m))          # only the lower triangle is the operand
    B = g.uniform(-1, 1, (m, n))
    C0 = g.uniform(-1, 1, (m, n))
    mine = oracle.dsymm(A, B, 2.0, -1.0, C0)
    ref = sb.dsymm(2.0, np.asfortranarray(A), np.asfortranarray(B), beta=-1.0,
                   c=np.asfortranarray(C0), side=0, lower=1)
    assert np.allclose(mine, ref, rtol=1e-13, atol=1e-13)


def test_dgemm_all_transposes():
    g = rng(5)
    A, B = g.standard_normal((7, 5)), g.standard_normal((5, 9))
    for ta in (0, 1):
        for tb in (0, 1):
            a = A.T.copy() if ta else A
            b = B.T.copy() if tb else B
            mine = oracle.dgemm(a, b, bool(ta), bool(tb))
            ref = sb.dgemm(1.0, np.asfortranarray(a), np.asfortranarray(b), trans_a=ta, trans_b=tb)
            assert np.allclose(mine, ref, rtol=1e-13, atol=1e-13)


# --------------------------------------------- plain-tier TRSM recursion emulation
def test_trsm_plain_emul_leaf_only_is_exact():
    # n < 2048: one chunk that is not a 2048-multiple -> a single exact leaf
    g = rng(9)
    n, m = 100, 13
    L = np.tril(g.uniform(-1, 1, (n, n))) + np.eye(n) * n
    B = g.uniform(-1, 1, (m, n))
    assert np.array_equal(oracle.tc_trsm_plain_emul(L, B, 256), oracle.dtrsm_rltn(L, B))


def test_trsm_plain_emul_fp16_exact_recursion():
    # L11 = L22 = I and L21 in {-1,0,1}: X1 = B1, X2 = B2 - B1 L21^T, every value an
    # integer below 2048, so FP16 rounding is exact and the recursion (top split at
    # 1024 plus the inner levels down to nb=256) must equal the exact solve.  A
    # transposed or shifted A21 would change X2.
    g = rng(10)
    n, m = 2048, 6
    L = np.eye(n)
    L[1024:, :1024] = g.integers(-1, 2, (1024, 1024)) * (g.random((1024, 1024)) < 0.004)
    L[256:512, :256] = g.integers(-1, 2, (256, 256)) * (g.random((256, 256)) < 0.01)
    B = g.integers(-3, 4, (m, n)).astype(np.float64)
    ex = oracle.dtrsm_rltn(L, B)
    assert np.abs(ex).max() < 2048
    assert np.array_equal(oracle.tc_trsm_plain_emul(L, B, 256), ex)


def test_trsm_plain_emul_error_model():
    # random data: the emulation differs from the exact solve by FP16 rounding of the
    # update operands, i.e. a relative error of order 2^-11, and not more
    g = rng(11)
    n, m = 2048, 8
    L = np.tril(g.uniform(-1, 1, (n, n)), -1) + np.diag(n + g.uniform(0, 1, n))
    B = g.uniform(-1, 1, (m, n))
    e = oracle.errors(oracle.tc_trsm_plain_emul(L, B, 256), oracle.dtrsm_rltn(L, B))
    assert 1e-7 < e["normwise"] < 2.0 ** -10


# ------------------------------------------------------------------ Ozaki
def test_ozaki_slice_reconstruction():
    g = rng(12)
    for phi in (0.5, 2.0):
        X = g.uniform(-0.5, 0.5, (9, 40)) * np.exp(phi * g.standard_normal((9, 40)))
        for S in (6, 8):
            sl, ex = oracle.ozaki_slice(X, S)
            assert sl.dtype == np.int8 and np.abs(sl.astype(int)).max() <= 127
            w = np.exp2(-7.0 * (np.arange(S) + 1))
            rec = np.einsum("s,sip->ip", w, sl.astype(np.float64)) * np.exp2(ex)[:, None]
            # truncation tail < 2^(e - 7S) per entry
            tol = np.exp2(ex - 7.0 * S)[:, None]
            assert np.all(np.abs(rec - X) < tol)
            mx = np.abs(X).max(axis=1)
            assert np.all((mx >= np.exp2(ex - 1)) & (mx < np.exp2(ex)))


def test_ozaki_slice_transposed_rows():
    g = rng(13)
    X = g.standard_normal((30, 7))
    a, ea = oracle.ozaki_slice(X, 7, trans=True)
    b, eb = oracle.ozaki_slice(X.T.copy(), 7, trans=False)
    assert np.array_equal(a, b) and np.array_equal(ea, eb)


def test_ozaki_gemm_exact_known_answer():
    # entries with at most 7*2 significant bits below the row max: A needs slices
    # 0..1, B slices 0..1, so with S = 3 no non-zero pair is dropped; every D_g is an
    # exact integer and the result equals the exact rational product up to the final
    # binary64 roundings of the three g-terms.
    g = rng(14)
    m, n, k = 5, 6, 50
    A = g.integers(-(1 << 13), 1 << 13, (m, k)) * np.exp2(g.integers(-30, 30, (m, 1)).astype(float))
    B = g.integers(-(1 << 13), 1 << 13, (n, k)) * np.exp2(g.integers(-30, 30, (n, 1)).astype(float))
    sa, ea = oracle.ozaki_slice(A, 3)
    sbs, eb = oracle.ozaki_slice(B, 3)
    assert not sa[2].any() and not sbs[2].any()
    C = oracle.ozaki_gemm(sa, ea, sbs, eb)
    for i in range(m):
        for j in range(n):
            exact = sum(Fraction(float(A[i, p])) * Fraction(float(B[j, p])) for p in range(k))
            assert abs(Fraction(C[i, j]) - exact) <= abs(exact) * Fraction(1, 2 ** 51) + Fraction(1, 2 ** 1000)


def test_ozaki_gemm_accuracy_vs_dgemm():
    g = rng(15)
    m, n, k = 24, 20, 300
    for phi in (0.5, 2.0):
        A = g.uniform(-0.5, 0.5, (m, k)) * np.exp(phi * g.standard_normal((m, k)))
        B = g.uniform(-0.5, 0.5, (n, k)) * np.exp(phi * g.standard_normal((n, k)))
        ref = oracle.dgemm(A, B, transb=True)
        prev = None
        for S in (4, 6, 8):
            sa, ea = oracle.ozaki_slice(A, S)
            sbs, eb = oracle.ozaki_slice(B, S)
            e = oracle.errors(oracle.ozaki_gemm(sa, ea, sbs, eb), ref)["maxrel"]
            if prev is not None:
                assert e < prev          # more slices, smaller error
            prev = e
        assert prev < 1e-13


def test_ozaki_kchunk_changes_only_rounding():
    g = rng(16)
    A, B = g.standard_normal((8, 200)), g.standard_normal((9, 200))
    sa, ea = oracle.ozaki_slice(A, 7)
    sbs, eb = oracle.ozaki_slice(B, 7)
    c1 = oracle.ozaki_gemm(sa, ea, sbs, eb)
    c2 = oracle.ozaki_gemm(sa, ea, sbs, eb, kchunk=64)
    assert oracle.errors(c2, c1)["maxrel"] < 1e-15


# ----------------------------------------------------------------- metrics
def test_error_metrics_known_values():
    R = np.array([[3.0, 0.0], [4.0, 1.0]])
    X = R + np.array([[0.0, 0.0], [0.0, 0.5]])
    e = oracle.errors(X, R)
    assert e["normwise"] == pytest.approx(0.5 / np.sqrt(26.0))
    assert e["maxrel"] == pytest.approx(0.5 / 4.0)
    S = np.ones((2, 2))
    assert oracle.errors(X, R, S)["componentwise"] == pytest.approx(0.5)
    # lower-only ignores the strictly upper part
    X2 = X.copy(); X2[0, 1] = 100.0
    assert oracle.errors(X2, R, lower=True)["maxrel"] == pytest.approx(0.5 / 4.0)


# ------------------------------------------------------------------ Cholesky
def _spd(g, n):
    B = g.standard_normal((n, n))
    return B @ B.T / n + np.eye(n)


def test_dpotrf_matches_numpy_cholesky():
    g = rng(31)
    for n in (1, 2, 7, 64, 129):
        A = _spd(g, n)
        Au = A.copy()
        Au[np.triu_indices(n, 1)] = np.nan          # the strict upper triangle is never read
        L, info = oracle.dpotrf_lower(Au)
        assert info == 0
        R = np.linalg.cholesky(A)
        assert np.abs(L - R).max() <= 1e-12 * np.abs(R).max()
        assert np.abs(L @ L.T - A).max() <= 1e-13 * np.abs(A).max() * n
        assert not np.triu(L, 1).any()


def test_dpotrf_info_reports_first_non_pd_minor():
    for n, k in ((5, 0), (6, 3), (9, 8)):
        A = np.eye(n)
        A[k, k] = -1.0
        L, info = oracle.dpotrf_lower(A)
        assert info == k + 1
        assert np.array_equal(np.diag(L)[:k], np.ones(k))      # steps before k completed
    v = np.arange(1.0, 5.0)
    assert oracle.dpotrf_lower(np.outer(v, v))[1] == 2          # rank one: the 2x2 minor is singular
    assert oracle.dpotrf_lower(np.array([[np.nan]]))[1] == 1


def test_ozaki_gemm_all_groups_is_exact_product_of_slices():
    """With every slice-pair group kept (ngroups = 2S - 1) the product is that of the
    sliced representations themselves, so it matches exact rational arithmetic on them
    up to the binary64 roundings of the group sums; keeping one group more than the
    triangular scheme (ngroups = S + 1) lies between the two."""
    g = rng(17)
    m, n, k, S = 4, 5, 40, 3
    A = g.standard_normal((m, k)) * np.exp2(g.integers(-20, 20, (m, 1)).astype(float))
    B = g.standard_normal((n, k)) * np.exp2(g.integers(-20, 20, (n, 1)).astype(float))
    sa, ea = oracle.ozaki_slice(A, S)
    sbs, eb = oracle.ozaki_slice(B, S)
    w = [Fraction(1, 2 ** (7 * (s + 1))) for s in range(S)]
    def rec(sl, ex, i, p):
        return sum(w[s] * int(sl[s, i, p]) for s in range(S)) * Fraction(2) ** int(ex[i])
    full = oracle.ozaki_gemm(sa, ea, sbs, eb, ngroups=2 * S - 1)
    tri = oracle.ozaki_gemm(sa, ea, sbs, eb)
    plus = oracle.ozaki_gemm(sa, ea, sbs, eb, ngroups=S + 1)
    for i in range(m):
        for j in range(n):
            exact = sum(rec(sa, ea, i, p) * rec(sbs, eb, j, p) for p in range(k))
            scale = sum(abs(rec(sa, ea, i, p) * rec(sbs, eb, j, p)) for p in range(k))
            assert abs(Fraction(full[i, j]) - exact) <= scale * Fraction(1, 2 ** 50)
            assert abs(plus[i, j] - float(exact)) <= abs(tri[i, j] - float(exact)) + float(scale) * 2 ** -50



# ------------------------------------------- extended-precision sampled columns
def _spread(g, shape, phi):
    return g.uniform(-0.5, 0.5, shape) * np.exp(phi * g.standard_normal(shape))


@pytest.mark.parametrize("phi", [0.5, 2.0])
def test_rowdot_xp_syr2k_columns_exact(phi):
    """SYR2K columns against exact rational arithmetic: after one rounding to FP64 the
    long-double sum is within 2^-53 |exact| + k 2^-62 sum|products| (products and
    sums carry a 64-bit significand).  A dropped second product, a swapped operand or a
    wrong column index is off by O(|C|)."""
    g = rng(int(phi * 10) + 3)
    n, k = 9, 37
    A, B = _spread(g, (n, k), phi), _spread(g, (n, k), phi)
    cols = [0, 4, 8, 2]
    out = oracle.rowdot_xp(A, B, cols, B, A)
    for c, j in enumerate(cols):
        for i in range(n):
            terms = [Fraction(float(A[i, p])) * Fraction(float(B[j, p])) +
                     Fraction(float(B[i, p])) * Fraction(float(A[j, p])) for p in range(k)]
            exact = sum(terms)
            mag = sum(abs(Fraction(float(A[i, p])) * Fraction(float(B[j, p]))) +
                      abs(Fraction(float(B[i, p])) * Fraction(float(A[j, p]))) for p in range(k))
            bound = abs(exact) * Fraction(1, 2 ** 53) + mag * k * Fraction(1, 2 ** 62)
            assert abs(Fraction(float(out[i, c])) - exact) <= bound, (i, j)


def test_rowdot_xp_gemm_column_and_fp64_agreement():
    """Column j of A @ B (X = A, Y = B^T) agrees with FP64 OpenBLAS to FP64 rounding,
    and the single-product form has no second term."""
    g = rng(5)
    m, k, n = 300, 257, 40
    A, B = g.uniform(-1, 1, (m, k)), g.uniform(-1, 1, (k, n))
    cols = [0, 17, 39]
    out = oracle.rowdot_xp(A, B.T, cols)
    R = (A @ B)[:, cols]
    assert np.max(np.abs(out - R)) <= 1e-13 * np.max(np.abs(R))
    # exactly representable case: small integers give the exact product
    Ai, Bi = g.integers(-50, 50, (m, k)).astype(np.float64), g.integers(-50, 50, (k, n)).astype(np.float64)
    assert np.array_equal(oracle.rowdot_xp(Ai, Bi.T, cols), (Ai @ Bi)[:, cols])
