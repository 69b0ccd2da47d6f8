"""Parity at the benchmark sizes, through the C ABI.

  * Recursive TRSM X*L^T = B at n = m = 16384, nb = 256 (six recursion levels, an
    8192-wide top update, 64 leaves; BLAS/tc_trsm.cu:31-137): split, FP32-FMA, plain and
    split + FP64-residual refinement, on 48 sampled rows of X against FP64 OpenBLAS
    dtrsm (scipy) of the same rows.  Tolerances (max-relative / normwise against the
    FP64 solve): split, FP32 and refined 5e-7 / 1e-7; plain 5e-5 / 5e-6; split within
    1.05x of the FP32-FMA tier in both metrics (the FP32-accuracy target), refinement
    no worse than the unrefined split solve.
  * Ozaki SYR2K C = A B^T + B A^T at n = 32768, k = 8192 (BLAS/tc_ozimmu_syr2k.cu:3-31)
    with entries U[-0.5,0.5]*exp(phi*N(0,1)), phi in {0.5, 2}: 64 sampled columns
    (lower part) against the oracle's extended-precision (long double) products.
    Configurations that meet max-relative error <= 1e-14 over the 64 columns:
    phi = 0.5: S = 8 + group g = 8, and S = 9; phi = 2: S = 9 + group g = 9, and S = 10.
    The others are held to the ceiling their dropped slice-pair groups set (measured on
    MI355X: phi = 0.5, S = 8: 1.2e-14; phi = 2, S = 8: 1.8e-12, S = 8+: 2.7e-14, S = 9:
    1.6e-14): 3e-14 for one dropped group fewer than the passing ones, 5e-12 for two.
"""
import numpy as np
import pytest

import oracle
from _tb import rng

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def h():
    import tensorblas
    hd = tensorblas.Handle()
    yield hd
    hd.close()


def _err(X, R):
    d = np.abs(X.astype(np.float64) - R)
    return {"maxrel": float(d.max() / np.abs(R).max()),
            "normwise": float(np.linalg.norm(d) / np.linalg.norm(R))}


@pytest.fixture(scope="module")
def trsm16384():
    import scipy.linalg.blas as sb
    n = m = 16384
    g = rng(0xC0F2)
    L = g.uniform(-1.0, 1.0, (n, n)).astype(np.float32)     # strict upper part never read
    L[np.diag_indices(n)] = (n + g.uniform(0.0, 1.0, n)).astype(np.float32)
    B = rng(0xC0F3).uniform(-1.0, 1.0, (m, n)).astype(np.float32)
    rows = np.sort(rng(7).choice(m, 48, replace=False))
    R = sb.dtrsm(1.0, np.asfortranarray(np.tril(L).astype(np.float64)),
                 np.asfortranarray(B[rows].astype(np.float64)), side=1, lower=1, trans_a=1)
    dL = torch.from_numpy(np.ascontiguousarray(L.T)).cuda()      # column-major, lda = n
    dB0 = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()     # column-major, ldb = m
    del L, B
    return dict(n=n, m=m, dL=dL, dB0=dB0, rows=rows, R=R, errs={})


TRSM_BOUNDS = {"split": (5e-7, 1e-7), "fp32": (5e-7, 1e-7), "refined": (5e-7, 1e-7),
               "plain": (5e-5, 5e-6)}


@pytest.mark.parametrize("tier", ["split", "fp32", "plain", "refined"])
def test_trsm_16384(h, trsm16384, tier):
    t = trsm16384
    n, m = t["n"], t["m"]
    dB = t["dB0"].clone()
    print(f"trsm 16384 {tier}", flush=True)
    if tier == "refined":
        h.trsm_refined("split", m, n, t["dL"].data_ptr(), n, dB.data_ptr(), m, nb=256)
    else:
        h.trsm(tier, m, n, t["dL"].data_ptr(), n, dB.data_ptr(), m, nb=256)
    h.synchronize()
    X = dB.cpu().numpy().T[t["rows"]]
    del dB
    e = _err(X, t["R"])
    t["errs"][tier] = e
    mx, nw = TRSM_BOUNDS[tier]
    assert np.isfinite(X).all()
    assert e["maxrel"] < mx and e["normwise"] < nw, (tier, e)


def test_trsm_16384_split_vs_fp32_and_refined(trsm16384):
    errs = trsm16384["errs"]
    if not all(t in errs for t in ("split", "fp32", "refined")):
        pytest.skip("needs the per-tier 16384 solves of this module")
    s, f, r = errs["split"], errs["fp32"], errs["refined"]
    assert s["maxrel"] <= 1.05 * f["maxrel"] and s["normwise"] <= 1.05 * f["normwise"], errs
    assert r["normwise"] <= s["normwise"] and r["maxrel"] <= s["maxrel"], errs


@pytest.fixture(scope="module", params=[0.5, 2.0], ids=["phi0.5", "phi2"])
def syr2k32768(request):
    phi = request.param
    n, k = 32768, 8192
    g = rng(0xC0F4 + int(phi * 10))
    A = g.uniform(-0.5, 0.5, (n, k)) * np.exp(phi * g.standard_normal((n, k)))
    B = g.uniform(-0.5, 0.5, (n, k)) * np.exp(phi * g.standard_normal((n, k)))
    cols = np.sort(rng(11).choice(n, 64, replace=False))
    print(f"syr2k 32768x8192 phi={phi}: extended-precision reference", flush=True)
    ref = oracle.rowdot_xp(A, B, cols, B, A)          # ref[:, c] = column cols[c], all rows
    dA = torch.from_numpy(A.T.copy()).cuda()           # column-major n x k (C order of A^T)
    dB = torch.from_numpy(B.T.copy()).cuda()
    del A, B
    dC = torch.zeros(n * n, device="cuda", dtype=torch.float64)
    yield dict(phi=phi, n=n, k=k, dA=dA, dB=dB, dC=dC, cols=cols, ref=ref)
    del dA, dB, dC
    torch.cuda.empty_cache()


# (phi, S, extra) -> max-relative bound over the 64 columns
OZAKI_BOUNDS = {(0.5, 8, 0): 3e-14, (0.5, 8, 1): 1e-14, (0.5, 9, 0): 1e-14,
                (2.0, 8, 0): 5e-12, (2.0, 8, 1): 3e-14, (2.0, 9, 0): 3e-14,
                (2.0, 9, 1): 1e-14, (2.0, 10, 0): 1e-14}


@pytest.mark.parametrize("S,extra", [(8, 0), (8, 1), (9, 0), (9, 1), (10, 0)])
def test_ozaki_syr2k_32768(h, syr2k32768, S, extra):
    t = syr2k32768
    phi, n, k = t["phi"], t["n"], t["k"]
    if (phi, S, extra) not in OZAKI_BOUNDS:
        pytest.skip("phi = 0.5 meets 1e-14 with fewer slice products")
    print(f"syr2k 32768x8192 phi={phi} S={S} extra={extra}", flush=True)
    h.set_ozaki_slices(S)
    h.set_ozaki_extra(extra)
    try:
        t["dC"].fill_(np.nan)
        h.syr2k("ozaki", n, k, 1.0, t["dA"].data_ptr(), n, t["dB"].data_ptr(), n, 0.0,
                t["dC"].data_ptr(), n, nb=256)
        h.synchronize()
    finally:
        h.set_ozaki_slices(7)
        h.set_ozaki_extra(0)
    C = t["dC"].view(n, n)             # row j of this view is column j of C
    worst = 0.0
    for c, j in enumerate(t["cols"]):
        got = C[j, j:].cpu().numpy()
        r = t["ref"][j:, c]
        assert np.isfinite(got).all()
        worst = max(worst, float(np.abs(got - r).max() / np.abs(r).max()))
    print(f"  maxrel over 64 columns: {worst:.3e}", flush=True)
    bound = OZAKI_BOUNDS[(phi, S, extra)]
    assert worst <= bound, (phi, S, extra, worst)
