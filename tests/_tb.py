"""Shared helpers of the test-suite: library loading and column-major device buffers."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_PY = os.path.join(ROOT, "tensor-blas_amd", "python")
LIB = os.path.join(ROOT, "tensor-blas_amd", "build", "libTensorBLAS.so")
HEADER = os.path.join(ROOT, "include", "TensorBLAS.h")
if PKG_PY not in sys.path:
    sys.path.insert(0, PKG_PY)

U32 = 2.0 ** -24
U64 = 2.0 ** -53


def gamma(k, u):
    """gamma_k = k*u/(1-k*u): the worst-case relative bound of a length-k dot product."""
    return k * u / (1.0 - k * u)


def rng(seed):
    return np.random.default_rng(seed)


def uniform(g, shape, lo=-1.0, hi=1.0, dtype=np.float32):
    return g.uniform(lo, hi, shape).astype(dtype)


def lower_tri_dominant(g, n, dtype=np.float32):
    """Strictly lower U[-1,1], diagonal n+U[0,1]; the TRSM benchmark's L."""
    L = np.tril(g.uniform(-1, 1, (n, n)), -1) + np.diag(n + g.uniform(0, 1, n))
    return L.astype(dtype)


def spread(g, shape, phi):
    """U[-0.5,0.5]*exp(phi*N(0,1)): the Ozaki benchmark's exponent spread."""
    return g.uniform(-0.5, 0.5, shape) * np.exp(phi * g.standard_normal(shape))


def to_dev(M, ld=None, fill=np.nan):
    """Column-major device copy of the 2-D array M with leading dimension ld (>= rows);
    padding rows hold `fill` so that reads outside the matrix show up."""
    import torch
    M = np.asarray(M)
    m, n = M.shape
    ld = max(1, m) if ld is None else ld
    buf = np.full((max(n, 1), ld), fill, dtype=M.dtype)
    buf[:n, :m] = M.T
    return torch.from_numpy(buf).cuda(), ld


def from_dev(t, m, n, ld):
    a = t.cpu().numpy().reshape(-1, ld)
    return np.ascontiguousarray(a[:n, :m].T)
