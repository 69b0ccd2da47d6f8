"""Batch plans (tb_batch_*): mixed SYRK / TRMM / TRSM instances in one persistent launch.

The contract is bit-equality with the single calls: the same schedule, exponents,
inverses and per-element MFMA sequence, whatever tile shapes and waves the plan chose
(from the tuning table, autotuned, or forced per bucket).  Accuracy against the FP64
oracle is then the single calls' (tests/test_gpu_blas3.py); a sample is re-checked here
as well."""
import numpy as np
import pytest

import oracle
from _tb import from_dev, lower_tri_dominant, rng, to_dev, uniform

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


SYRKS = [(1, 1), (7, 33), (64, 64), (130, 70), (300, 300), (513, 200)]      # (n, k)
TRSMS = [(1, 1), (64, 64), (100, 300), (257, 1000), (65, 2048), (300, 513), (70, 3001)]  # (m, n)
TRMMS = [(1, 1), (5, 77), (64, 64), (129, 40), (300, 257), (700, 130)]       # (m, n)


def make_problems(seed):
    g = rng(seed)
    probs = []
    for n, k in SYRKS:
        A = uniform(g, (n, k))
        C0 = uniform(g, (n, n))
        probs.append(("syrk", n, k, A, C0, 1.0 if n % 2 else -0.5, 0.0 if n % 3 else 0.75))
    for m, n in TRSMS:
        L = lower_tri_dominant(g, n)
        B = uniform(g, (m, n))
        probs.append(("trsm", m, n, L, B))
    for m, n in TRMMS:
        L = uniform(g, (m, m))
        L[np.triu_indices(m, 1)] = np.nan          # the strict upper triangle is never read
        B = uniform(g, (m, n))
        C0 = uniform(g, (m, n))
        probs.append(("trmm", m, n, L, B, C0, 1.0 if m % 2 else -1.5))
    order = np.random.default_rng(seed).permutation(len(probs))
    return [probs[i] for i in order]


def upload(probs):
    """device buffers (ld padded) and batch entries"""
    dev, entries = [], []
    for p in probs:
        if p[0] == "syrk":
            _, n, k, A, C0, alpha, beta = p
            dA, lda = to_dev(A, n + 3)
            dC, ldc = to_dev(C0, n + 1)
            dev.append((dA, dC))
            entries.append(dict(op="syrk", n=n, k=k, A=dA.data_ptr(), lda=lda, C=dC.data_ptr(),
                                ldc=ldc, alpha=alpha, beta=beta))
        elif p[0] == "trmm":
            _, m, n, L, B, C0, alpha = p
            dL, lda = to_dev(L, m + 1)
            dB, ldb = to_dev(B, m + 4)
            dC, ldc = to_dev(C0, m + 2)
            dev.append((dL, dB, dC))
            entries.append(dict(op="trmm", m=m, n=n, A=dL.data_ptr(), lda=lda, B=dB.data_ptr(),
                                ldb=ldb, C=dC.data_ptr(), ldc=ldc, alpha=alpha))
        else:
            _, m, n, L, B = p
            dL, lda = to_dev(L)
            dB, ldb = to_dev(B, m + 2)
            dev.append((dL, dB))
            entries.append(dict(op="trsm", m=m, n=n, A=dL.data_ptr(), lda=lda, C=dB.data_ptr(),
                                ldc=ldb))
    return dev, entries


def single_results(h, tier, probs):
    out = []
    for p in probs:
        if p[0] == "syrk":
            _, n, k, A, C0, alpha, beta = p
            dA, lda = to_dev(A, n + 3)
            dC, ldc = to_dev(C0, n + 1)
            h.syrk(tier, n, k, alpha, dA.data_ptr(), lda, beta, dC.data_ptr(), ldc)
            h.synchronize()
            out.append(from_dev(dC, n, n, ldc))
        elif p[0] == "trmm":
            _, m, n, L, B, C0, alpha = p
            dL, lda = to_dev(L, m + 1)
            dB, ldb = to_dev(B, m + 4)
            dC, ldc = to_dev(C0, m + 2)
            h.trmm(tier, m, n, alpha, dL.data_ptr(), lda, dB.data_ptr(), ldb, dC.data_ptr(), ldc)
            h.synchronize()
            out.append(from_dev(dC, m, n, ldc))
        else:
            _, m, n, L, B = p
            dL, lda = to_dev(L)
            dB, ldb = to_dev(B, m + 2)
            h.trsm(tier, m, n, dL.data_ptr(), lda, dB.data_ptr(), ldb, nb=256)
            h.synchronize()
            out.append(from_dev(dB, m, n, ldb))
    return out


def batch_outputs(probs, dev, entries):
    res = []
    for p, d, e in zip(probs, dev, entries):
        rows = p[1]
        cols = p[1] if p[0] == "syrk" else p[2]
        res.append(from_dev(d[-1], rows, cols, e["ldc"]))
    return res


def check_equal(probs, ref, got):
    for p, r, g in zip(probs, ref, got):
        if p[0] == "syrk":
            assert np.array_equal(np.tril(g), np.tril(r)), p[:3]
            assert np.array_equal(np.triu(g, 1), np.triu(p[4], 1))     # upper untouched
        else:
            assert np.array_equal(g, r), p[:3]


@pytest.mark.parametrize("tier", ["split", "plain"])
@pytest.mark.parametrize("autotune", [False, True])
def test_batch_equals_single_calls(h, tbm, tier, autotune):
    probs = make_problems(0xC0F6)
    ref = single_results(h, tier, probs)
    dev, entries = upload(probs)
    b = tbm.Batch(h, tier, entries, nb=256, autotune=autotune)
    info = b.info()
    assert info["instances"] == len(probs) and info["items"] > 0
    assert (info["tune_us"] > 0) == autotune
    tiles = b.tiles()
    assert tiles["waves"] in ((4, 8) if tier == "split" else (4,))
    assert set(tiles["syrk"] + tiles["trmm"]) <= {0, 1}
    b.run()
    h.synchronize()
    check_equal(probs, ref, batch_outputs(probs, dev, entries))
    b.close()


@pytest.mark.parametrize("tier", ["split", "plain"])
@pytest.mark.parametrize("tiles", ["panel", "square"])
@pytest.mark.parametrize("waves", [4, 8])
def test_batch_every_tile_choice_equals_single_calls(h, tbm, tier, tiles, waves):
    """Each forced tile shape (64 x 256 or 128 x 128 for every SYRK / TRMM instance) and
    wave count (8 only for 2-plane plans) gives the single calls' bits."""
    if tier == "plain" and waves == 8:
        pytest.skip("1-plane panels keep four waves")
    probs = make_problems(0xC0F7)
    ref = single_results(h, tier, probs)
    dev, entries = upload(probs)
    b = tbm.Batch(h, tier, entries, tiles=tiles, waves=waves)
    t = b.tiles()
    assert t["waves"] == waves
    assert set(t["syrk"] + t["trmm"]) == {1 if tiles == "square" else 0}
    b.run()
    h.synchronize()
    check_equal(probs, ref, batch_outputs(probs, dev, entries))
    b.close()


def test_batch_accuracy_sample(h, tbm):
    probs = make_problems(11)
    dev, entries = upload(probs)
    b = tbm.Batch(h, "split", entries)
    b.run()
    h.synchronize()
    for p, g in zip(probs, batch_outputs(probs, dev, entries)):
        if p[0] == "trsm":
            _, m, n, L, B = p
            e = oracle.errors(g, oracle.dtrsm_rltn(L, B))
            assert e["normwise"] < 1e-6, (p[:3], e)
        elif p[0] == "trmm":
            _, m, n, L, B, C0, alpha = p
            e = oracle.errors(g, oracle.dtrmm(np.tril(np.nan_to_num(L)), B, alpha))
            assert e["normwise"] < 1e-6, (p[:3], e)
        else:
            _, n, k, A, C0, alpha, beta = p
            R = oracle.dsyrk(A, alpha, beta, C0)
            e = oracle.errors(np.tril(g), np.tril(R))
            assert e["normwise"] < 1e-6, (p[:3], e)
    b.close()


def test_batch_graph_replays_match_eager(h, tbm):
    """The captured graph, replayed with eager work in between, gives the eager result
    on every replay (inputs restored before each run)."""
    probs = make_problems(5)
    dev, entries = upload(probs)
    saved = [d[-1].clone() for d in dev]
    eager = tbm.Batch(h, "split", entries, graph=False)
    eager.run()
    h.synchronize()
    want = batch_outputs(probs, dev, entries)
    eager.close()
    g = tbm.Batch(h, "split", entries, graph=True)
    for rep in range(3):
        for d, s in zip(dev, saved):
            d[-1].copy_(s)
        torch.cuda.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()     # unrelated eager work
        g.run()
        h.synchronize()
        got = batch_outputs(probs, dev, entries)
        for p, w, x in zip(probs, want, got):
            assert np.array_equal(x, w), (rep, p[:3])
    g.close()


def test_batch_empty_and_degenerate(h, tbm):
    b = tbm.Batch(h, "split", [])
    b.run()
    h.synchronize()
    b.close()
    z = torch.zeros(4, device="cuda")
    b = tbm.Batch(h, "split", [dict(op="trsm", m=0, n=3, A=z.data_ptr(), lda=3, C=z.data_ptr(), ldc=1),
                               dict(op="syrk", n=0, k=5, A=z.data_ptr(), lda=1, C=z.data_ptr(), ldc=1)])
    b.run()
    h.synchronize()
    b.close()


def test_batch_rejects_bad_entries(h, tbm):
    z = torch.zeros(16, device="cuda")
    bad = [dict(op="trsm", m=4, n=4, A=z.data_ptr(), lda=3, C=z.data_ptr(), ldc=4),    # lda < n
           dict(op="syrk", n=4, k=4, A=0, lda=4, C=z.data_ptr(), ldc=4)]              # null A
    for e in bad:
        with pytest.raises(tbm.TensorBLASError):
            tbm.Batch(h, "split", [e])
    with pytest.raises(tbm.TensorBLASError):
        tbm.Batch(h, "fp32", [])          # FP16 tiers only
    trmm_bad = [dict(op="trmm", m=4, n=4, A=z.data_ptr(), lda=4, B=0, ldb=4, C=z.data_ptr(), ldc=4),
                dict(op="trmm", m=4, n=2, A=z.data_ptr(), lda=4, B=z.data_ptr(), ldb=3, C=z.data_ptr(),
                     ldc=4)]                                                          # null B; ldb < m
    for e in trmm_bad:
        with pytest.raises(tbm.TensorBLASError):
            tbm.Batch(h, "split", [e])
    with pytest.raises(tbm.TensorBLASError):
        tbm.Batch(h, "plain", [], waves=8)  # 1-plane panels keep four waves
