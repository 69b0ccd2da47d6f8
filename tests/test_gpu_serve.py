"""The request loop (tb_serve): mixed routines and tiers on concurrent slots.

Each request must compute exactly what it computes alone (bitwise: the kernels are
deterministic and slots share nothing but the device); admission must respect the
HBM budget and first-come first-served order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def tbm():
    import tensorblas
    return tensorblas


def make_requests(seed, sizes=(96, 200, 320)):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    reqs, keep = [], []
    ops = ["syrk", "syr2k", "trmm", "trsm", "symm"]
    q = 0
    for tier in ("plain", "split", "ozaki"):
        dt = torch.float64 if tier == "ozaki" else torch.float32
        for op in ops:
            n = sizes[q % len(sizes)]
            q += 1

            def rnd(*shape):
                return (torch.rand(*shape, device="cuda", dtype=dt, generator=g) * 2 - 1).contiguous()
            r = dict(op=op, tier=tier, nb=256, alpha=1.0, beta=0.0)
            if op in ("syrk", "syr2k"):
                A, B, C = rnd(n, n), rnd(n, n), rnd(n, n)     # column-major n x k (k = n)
                r.update(n=n, k=n, A=A.data_ptr(), lda=n, B=B.data_ptr(), ldb=n, C=C.data_ptr(),
                         ldc=n, beta=0.5)
            elif op == "trsm":
                L, B = rnd(n, n), rnd(n, n)
                L.diagonal().copy_(n + torch.rand(n, device="cuda", dtype=dt, generator=g))
                A, C = L, B
                r.update(m=n, n=n, A=L.data_ptr(), lda=n, C=B.data_ptr(), ldc=n)
            else:
                A, B, C = rnd(n, n), rnd(n, n), rnd(n, n)
                r.update(m=n, n=n, A=A.data_ptr(), lda=n, B=B.data_ptr(), ldb=n, C=C.data_ptr(),
                         ldc=n, beta=0.25 if op == "symm" else 0.0)
            reqs.append(r)
            keep.append((A, B, C, C.clone()))       # operands and the output's initial value
    return reqs, keep


def restore(keep):
    for _, _, C, C0 in keep:
        C.copy_(C0)
    torch.cuda.synchronize()


def slots(tbm, n):
    hs = []
    for _ in range(n):
        h = tbm.Handle(torch.cuda.Stream())
        h.reserve_workspace(64 << 20, 64 << 20)
        hs.append(h)
    return hs


def test_serve_mix_equals_alone(tbm):
    reqs, keep = make_requests(0xC0F7)
    alone = []
    h1 = slots(tbm, 1)
    for r, (_, _, C, _) in zip(reqs, keep):
        res = tbm.serve(h1, 1 << 34, [dict(r, arrival_us=0.0)])
        assert res[0]["status"] == 0 and res[0]["done_us"] > 0, r
        alone.append(C.clone())
    restore(keep)
    hs = slots(tbm, 4)
    # all at once (maximal overlap), then staggered arrivals
    for spacing in (0.0, 50.0):
        restore(keep)
        mix = [dict(r, arrival_us=spacing * i) for i, r in enumerate(reqs)]
        res = tbm.serve(hs, 1 << 34, mix)
        for i, (r, x) in enumerate(zip(res, reqs)):
            assert r["status"] == 0 and r["done_us"] >= r["admit_us"] - 50.0, (i, r)
            assert torch.equal(keep[i][2], alone[i]), (spacing, x["op"], x["tier"], x["n"])
        assert len({r["slot"] for r in res}) > 1           # the work really was spread


def test_serve_budget_and_order(tbm):
    reqs, keep = make_requests(3, sizes=(128,))
    hs = slots(tbm, 4)
    # a budget that holds one request at a time: strictly sequential, in arrival order
    one = [dict(r, bytes=1000, arrival_us=0.0) for r in reqs]
    res = tbm.serve(hs, 1500, one)
    assert all(r["status"] == 0 for r in res)
    for a, b in zip(res, res[1:]):
        assert b["admit_us"] >= a["admit_us"]
        assert b["admit_us"] + 1.0 >= a["done_us"] - 200.0     # device vs host clock slack
    # a request larger than the budget is refused, the others run
    big = [dict(r, bytes=1000, arrival_us=0.0) for r in reqs[:3]]
    big[1]["bytes"] = 10 ** 9
    res = tbm.serve(hs, 5000, big)
    assert [r["status"] for r in res] == [0, 3, 0]


def test_serve_rejects_bad_arguments(tbm):
    hs = slots(tbm, 1)
    with pytest.raises(tbm.TensorBLASError):
        tbm.serve(hs, 0, [])
    z = torch.zeros(4, device="cuda")
    with pytest.raises(tbm.TensorBLASError):
        tbm.serve(hs, 1 << 20, [dict(op="syrk", tier="plain", n=1, k=1, A=z.data_ptr(), lda=1,
                                     C=z.data_ptr(), ldc=1, arrival_us=-1.0)])
    with pytest.raises(tbm.TensorBLASError):
        tbm.serve(hs, 1 << 20, [], reserve_cus=32, partition=3)
    with pytest.raises(tbm.TensorBLASError):          # every slot kept for small requests
        tbm.serve(hs, 1 << 20, [], small_us=100.0, small_slots=1)
    with pytest.raises(tbm.TensorBLASError):
        tbm.serve(hs, 1 << 20, [], small_us=-1.0)


@pytest.mark.parametrize("reserve_cus,partition", [(0, 0), (32, 0), (32, 1), (32, 2)])
def test_serve_lanes(tbm, reserve_cus, partition):
    """Requests estimated above large_us run at most max_large at a time; small ones
    keep flowing (with reserve_cus, on CUs the large requests leave free: partition 0
    caps their GEMMs, 1 runs them on a CU-masked stream; partition 2 runs the small ones
    on CU-masked streams instead); every result still equals the isolated one."""
    reqs, keep = make_requests(9, sizes=(128, 448))
    alone, est = [], []
    h1 = slots(tbm, 1)
    for r, (_, _, C, _) in zip(reqs, keep):
        est.append(tbm.serve(h1, 1 << 34, [dict(r, arrival_us=0.0)])[0]["est_us"])
        alone.append(C.clone())
    restore(keep)
    thr = float(np.median(est))                # the library's own estimates split the set
    # partition 2 masks a stream per slot: each takes a hardware queue, of which a
    # process has 4 unless GPU_MAX_HW_QUEUES says otherwise
    hs = slots(tbm, 3 if partition == 2 else 6)
    res = tbm.serve(hs, 1 << 34, [dict(r, arrival_us=0.0) for r in reqs], large_us=thr, max_large=1,
                    reserve_cus=reserve_cus, partition=partition)
    large = [x for x in res if x["est_us"] > thr]
    assert 0 < len(large) < len(res)
    large.sort(key=lambda x: x["admit_us"])
    for a, b in zip(large, large[1:]):
        assert b["admit_us"] + 200.0 >= a["done_us"]          # never two large at once
    for i, x in enumerate(res):
        assert x["status"] == 0
        assert torch.equal(keep[i][2], alone[i]), (reqs[i]["op"], reqs[i]["tier"], reqs[i]["n"])


def test_serve_shortest_first(tbm):
    """With one slot and every request waiting from the start, TB_SERVE_SHORTEST admits
    in increasing estimated time, FIFO in arrival order; the results do not change."""
    reqs, keep = make_requests(21, sizes=(64, 384, 160, 256))
    hs = slots(tbm, 1)
    mix = [dict(r, arrival_us=0.0) for r in reqs]
    res = tbm.serve(hs, 1 << 34, mix, order="shortest")
    assert all(x["status"] == 0 for x in res)
    by_admit = sorted(range(len(res)), key=lambda i: res[i]["admit_us"])
    ests = [res[i]["est_us"] for i in by_admit]
    assert ests == sorted(ests)
    first = [C.clone() for _, _, C, _ in keep]
    restore(keep)
    res = tbm.serve(hs, 1 << 34, mix, order="fifo")
    by_admit = sorted(range(len(res)), key=lambda i: (res[i]["admit_us"], i))
    assert by_admit == list(range(len(res)))
    for i, (_, _, C, _) in enumerate(keep):
        assert torch.equal(C, first[i]), (reqs[i]["op"], reqs[i]["tier"])


@pytest.mark.parametrize("order,reserve_cus,partition",
                         [("fifo", 0, 0), ("shortest", 0, 0), ("shortest-small", 0, 0), ("shortest", 32, 0),
                          ("shortest", 32, 1), ("shortest", 32, 2)])
def test_serve_small_lane(tbm, order, reserve_cus, partition):
    """With small_slots slots kept for requests estimated below small_us, the other
    requests never hold more than nslots - small_slots slots at once, small ones are
    admitted before the others that wait with them (with reserve_cus, the others leave
    CUs free for them, or run beside them on masked streams: partition 0, 1, 2), and
    every result still equals the isolated one."""
    reqs, keep = make_requests(17, sizes=(96, 384, 160))
    alone, est = [], []
    h1 = slots(tbm, 1)
    for r, (_, _, C, _) in zip(reqs, keep):
        est.append(tbm.serve(h1, 1 << 34, [dict(r, arrival_us=0.0)])[0]["est_us"])
        alone.append(C.clone())
    restore(keep)
    small = float(np.percentile(est, 40))
    large = float(np.percentile(est, 85))
    hs = slots(tbm, 3)
    res = tbm.serve(hs, 1 << 34, [dict(r, arrival_us=0.0) for r in reqs], large_us=large,
                    max_large=1, order=order, small_us=small, small_slots=1,
                    reserve_cus=reserve_cus, partition=partition)
    assert all(x["status"] == 0 for x in res)
    regular = [x for x in res if x["est_us"] >= small]
    smalls = [x for x in res if x["est_us"] < small]
    assert 0 < len(smalls) < len(res) and len(regular) > 2
    for x in regular:            # at most 3 - 1 = 2 of them at any admission
        busy = [y for y in regular if y is not x and y["admit_us"] <= x["admit_us"]
                and y["done_us"] > x["admit_us"] + 200.0]
        assert len(busy) <= 1, (x, busy)
    # everything is waiting from t = 0 and the small lane is served first
    assert max(x["admit_us"] for x in smalls) <= min(x["admit_us"] for x in regular)
    for i, x in enumerate(res):
        assert torch.equal(keep[i][2], alone[i]), (reqs[i]["op"], reqs[i]["tier"], reqs[i]["n"])
