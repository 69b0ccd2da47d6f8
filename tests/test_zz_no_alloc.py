"""The no-allocation entry point tb_syrk_noalloc (no reference counterpart; for callers
that cannot allocate device memory): through the C ABI it must give tb_syrk's bits in
every tier it offers, on two calls in a row (the second smaller, over the scratch the
first left behind), allocate nothing (the handle's workspace stays empty), refuse a
problem larger than its fixed scratch with TB_ERR_SCRATCH before launching, and refuse
the Ozaki tier."""
import ctypes

import numpy as np
import pytest

from _tb import from_dev, rng, to_dev, uniform

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TB_ERR_INVALID, TB_ERR_SCRATCH = 1, 4


def _syrk_pair(L, h_na, h_ref, tier, n, k, alpha, beta, g):
    A = uniform(g, (n, k))
    C0 = uniform(g, (n, n))
    dA, lda = to_dev(A, n + 3)
    dC1, ldc = to_dev(C0, n + 2)
    dC2, _ = to_dev(C0, n + 2)
    t = {"plain": 0, "split": 1, "fp32": 3}[tier]
    st1 = L.tb_syrk_noalloc(h_na.ptr, t, n, k, alpha, dA.data_ptr(), lda, beta, dC1.data_ptr(), ldc, 256)
    st2 = L.tb_syrk(h_ref.ptr, t, n, k, alpha, dA.data_ptr(), lda, beta, dC2.data_ptr(), ldc, 256)
    h_na.synchronize()
    h_ref.synchronize()
    assert (st1, st2) == (0, 0)
    c1 = dC1.cpu().numpy()
    c2 = dC2.cpu().numpy()
    # the whole buffers, padding and upper triangle included, are identical
    assert np.array_equal(c1.view(np.uint32), c2.view(np.uint32)), (tier, n, k)
    # and the result is a SYRK (a transposed or misplaced operand is O(1) off)
    R = alpha * (A.astype(np.float64) @ A.T.astype(np.float64)) + beta * C0
    got = from_dev(dC1, n, n, ldc)
    lo = np.tril_indices(n)
    assert np.abs(got[lo] - R[lo]).max() <= 5e-3 * np.abs(R[lo]).max()


@pytest.mark.parametrize("tier", ["split", "plain", "fp32"])
def test_noalloc_equals_syrk_twice_second_smaller(tier):
    import tensorblas
    L = tensorblas.lib()
    h_na = tensorblas.Handle()
    h_ref = tensorblas.Handle()
    try:
        g = rng(0x7A + len(tier))
        _syrk_pair(L, h_na, h_ref, tier, 1031, 1300, 1.25, 0.5, g)    # 2 planes: 5.5 MB of scratch
        _syrk_pair(L, h_na, h_ref, tier, 301, 77, -0.75, 0.0, g)      # smaller, on old contents
        assert h_na.workspace_bytes() == 0                             # nothing was allocated
    finally:
        h_na.close()
        h_ref.close()


def test_noalloc_rejects_oversize_and_ozaki():
    import tensorblas
    L = tensorblas.lib()
    h = tensorblas.Handle()
    try:
        # split planes of 4096 x 8192 need 4 * 4096 * 8192 B = 128 MiB > 64 MiB
        n, k = 4096, 8192
        dA = torch.zeros(n * k, device="cuda")
        dC = torch.full((n * n,), 3.0, device="cuda")
        assert L.tb_syrk_noalloc(h.ptr, 1, n, k, 1.0, dA.data_ptr(), n, 0.0, dC.data_ptr(), n,
                                 256) == TB_ERR_SCRATCH
        h.synchronize()
        assert bool((dC == 3.0).all())                    # nothing was launched
        d64 = torch.zeros(64, dtype=torch.float64, device="cuda")
        assert L.tb_syrk_noalloc(h.ptr, 2, 4, 4, 1.0, d64.data_ptr(), 4, 0.0, d64.data_ptr(), 4,
                                 256) == TB_ERR_INVALID
        assert L.tb_status_string(TB_ERR_SCRATCH) == b"problem exceeds the no-allocation scratch"
        assert h.workspace_bytes() == 0
    finally:
        h.close()
