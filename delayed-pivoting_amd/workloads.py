"""Synthetic inputs of the delayed-pivoting LU workloads (seeded, reproducible).

Matrices are returned as Fortran-ordered float64 numpy arrays (column-major, lda = m).
  normal(m, n, seed)          entries iid N(0,1)
  delay_stress(n, seed)       N(0,1) with 5% exactly-zero columns, a disjoint 5% of columns
                              scaled by 1e-14 and one 64x64 aligned diagonal block of rank 32
  diag_rank(n, r)             zero except diag(2..r+1) = r..1 (the reference example's matrix,
                              delayed_lu/example.cpp:215-220)
"""
from __future__ import annotations

import numpy as np


def normal(m: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((n, m)).T)


def delay_stress(n: int, seed: int, frac_zero: float = 0.05, frac_tiny: float = 0.05,
                 block: int = 64, block_rank: int = 32):
    """Returns (matrix, zero_cols, tiny_cols, block_offset)."""
    rng = np.random.default_rng(seed)
    a = np.asfortranarray(rng.standard_normal((n, n)).T)
    z = int(round(frac_zero * n))
    t = int(round(frac_tiny * n))
    nblocks = n // block
    off = int(rng.integers(0, nblocks)) * block
    block_cols = np.arange(off, off + block)
    others = np.setdiff1d(np.arange(n), block_cols)
    perm = rng.permutation(others)
    zero_cols = np.sort(perm[:z])
    tiny_cols = np.sort(perm[z:z + t])
    x = rng.standard_normal((block, block_rank))
    y = rng.standard_normal((block, block_rank))
    a[off:off + block, off:off + block] = x @ y.T
    a[:, tiny_cols] *= 1e-14
    a[:, zero_cols] = 0.0
    return a, zero_cols, tiny_cols, off


def conditioned(n: int, seed: int, smin: float = 1e-6) -> np.ndarray:
    """A = U diag(s) V^T with U, V the Q factors of Householder QR of seeded Gaussians and
    s_i = smin ** (i / (n - 1)) (geometric from 1 to smin): the mixed-precision workload."""
    rng = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rng.standard_normal((n, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = smin ** (np.arange(n) / max(n - 1, 1))
    return np.asfortranarray((u * s) @ v.T)


def diag_rank(n: int, r: int) -> np.ndarray:
    a = np.zeros((n, n), order="F")
    for k in range(r):
        a[k + 2, k + 2] = float(r - k)
    return a
