"""The bench's reference for the Ozaki configurations (bench.exact_rowdot: exact integer
slice products on FP64 BLAS) against exact rational arithmetic and against the oracle's
extended-precision products, so that the errors the bench reports for the 1e-14 target
are not those of its reference."""
from fractions import Fraction

import numpy as np
import pytest

import bench
import oracle


@pytest.mark.parametrize("phi", [0.5, 2.0])
def test_exact_rowdot_syr2k_columns(phi):
    g = np.random.default_rng(int(phi * 100) + 1)
    n, k = 64, 300
    A = g.uniform(-0.5, 0.5, (n, k)) * np.exp(phi * g.standard_normal((n, k)))
    B = g.uniform(-0.5, 0.5, (n, k)) * np.exp(phi * g.standard_normal((n, k)))
    cols = np.array([0, 63, 7, 31])
    got = bench.exact_rowdot([(A, B), (B, A)], cols)
    xp = oracle.rowdot_xp(A, B, cols, B, A)
    assert np.abs(got - xp).max() <= 4e-16 * np.abs(xp).max()
    for i, c in [(0, 0), (40, 1), (63, 3), (12, 2)]:
        j = cols[c]
        exact = sum(Fraction(float(A[i, p])) * Fraction(float(B[j, p])) +
                    Fraction(float(B[i, p])) * Fraction(float(A[j, p])) for p in range(k))
        # three 20-bit slices keep 60 bits of each row: the truncation is far below 2^-53
        assert abs(Fraction(float(got[i, c])) - exact) <= abs(exact) * Fraction(1, 2 ** 51)


def test_exact_rowdot_long_k_is_exact_on_integers():
    """k = 32768 (SYMM's inner dimension) uses 19-bit slices; integer inputs whose
    products overflow FP64's exact range in a plain sum still come out exact."""
    g = np.random.default_rng(9)
    k = 32768
    X = g.integers(-2 ** 26, 2 ** 26, (5, k)).astype(np.float64)
    Y = g.integers(-2 ** 26, 2 ** 26, (3, k)).astype(np.float64)
    got = bench.exact_rowdot([(X, Y)], np.array([2, 0]))
    for i in range(5):
        for c, j in enumerate([2, 0]):
            exact = sum(int(a) * int(b) for a, b in zip(X[i].astype(np.int64), Y[j].astype(np.int64)))
            assert abs(int(got[i, c]) - exact) <= abs(exact) * 2 ** -52
