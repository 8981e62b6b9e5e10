"""Panel pairs (two full panels sharing one K = 2 nb far update, with the second panel's U block
row formed from the first's) against factorizations without pairs, at nb 64 and 128, FP64 and
FP32.

merge_rows = 1 pairs every eligible pair of full panels, merge_rows = -1 none.  The pivots must
agree exactly, the factors to rounding (the pair's far update sums K = 2 nb in one GEMM), and
the paired factorization must meet the oracle bar of the other GPU tests.
"""
import numpy as np
import pytest

from helpers import dlu, oracle, workloads
from test_gpu_getrf import FACTOR_RTOL, check_parity

EPS32_OVER_EPS64 = float(np.finfo(np.float32).eps / np.finfo(np.float64).eps)

pytestmark = pytest.mark.gpu


def _factor(a, merge_rows, nb, dtype=np.float64):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    m, n = a.shape
    t, lda = dlu.to_device(a)
    options = {"nb": nb, "merge_rows": merge_rows}
    if dtype == np.float32:
        t = t.to(torch.float32)
        out = dlu.sgetrf_device(m, n, t.data_ptr(), lda, options=options, margins=True)
    else:
        out = dlu.getrf_device(m, n, t.data_ptr(), lda, options=options, margins=True)
    torch.cuda.synchronize()
    st, ipiv, jpiv, info, rep, marg = out
    assert st == dlu.Status.SUCCESS, dlu.status_string(st)
    return np.asfortranarray(t.double().cpu().numpy()[:, :m].T), ipiv, jpiv, info, marg


@pytest.mark.parametrize("nb", [64, 128])
@pytest.mark.parametrize("shape", [(1300, 1300), (2100, 900), (700, 1500)])
def test_pairs_match_single_panels_fp64(nb, shape):
    a = workloads.normal(shape[0], shape[1], 31 + nb + shape[0])
    p = _factor(a, 1, nb)
    s = _factor(a, -1, nb)
    assert p[3] == s[3] and np.array_equal(p[1], s[1]) and np.array_equal(p[2], s[2])
    assert np.linalg.norm(p[0] - s[0]) <= FACTOR_RTOL * np.linalg.norm(s[0])
    check_parity(a, p[0], p[1], p[2], p[3], p[4])


@pytest.mark.parametrize("nb", [64, 128])
def test_pairs_match_single_panels_fp32(nb):
    a = workloads.normal(1500, 1500, 77).astype(np.float32).astype(np.float64)
    p = _factor(a, 1, nb, np.float32)
    s = _factor(a, -1, nb, np.float32)
    assert p[3] == s[3] and np.array_equal(p[1], s[1]) and np.array_equal(p[2], s[2])
    # FP32 factors of two summation orders differ by far more than eps32 in the late Schur
    # complements; both must be backward stable: ||PAQ - LU||_F / (n eps32 ||A||_F) <= 10.
    for f in (p, s):
        assert oracle.backward_error(a, f[0], f[1], f[2], f[3]) / EPS32_OVER_EPS64 <= 10.0


def test_pairs_with_zero_columns_and_stress():
    # Zero columns end pairs half-way (the first panel's update then runs alone); the stress
    # matrix defers its tail.
    a = workloads.normal(1400, 1400, 5)
    a[:, [200, 455, 900]] = 0.0
    p = _factor(a, 1, 64)
    check_parity(a, p[0], p[1], p[2], p[3], p[4])
    b = workloads.delay_stress(1024, 9)[0]
    p = _factor(b, 1, 64)
    check_parity(b, p[0], p[1], p[2], p[3], p[4])
