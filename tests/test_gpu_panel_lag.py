"""The one-exchange-per-step panel kernel against the exchange-and-move kernel and the oracle.

Both panel kernels run the same multiply and fma sequence on every element (multipliers by the
reciprocal of the pivot, one rank-1 update per step inside the panel), so their factorizations
must agree bit for bit: factors, ipiv, jpiv and info.  The one-exchange kernel is also held to
the oracle bar of the other GPU tests (pivots bit-exact against delayed_getrf with nb = 64).
"""
import numpy as np
import pytest

from helpers import dlu, oracle, workloads
from test_gpu_getrf import check_parity

pytestmark = pytest.mark.gpu


def _factor(a, kernel, dtype=np.float64, **opts):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    m, n = a.shape
    t, lda = dlu.to_device(a)
    options = dict(opts, panel_kernel=kernel)
    if dtype == np.float32:
        t = t.to(torch.float32)
        st, ipiv, jpiv, info, rep, marg = dlu.sgetrf_device(m, n, t.data_ptr(), lda,
                                                            options=options, margins=True)
    else:
        st, ipiv, jpiv, info, rep, marg = dlu.getrf_device(m, n, t.data_ptr(), lda,
                                                           options=options, margins=True)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS, dlu.status_string(st)
    lu = np.asfortranarray(t.double().cpu().numpy()[:, :m].T)
    return lu, ipiv, jpiv, info, rep, marg


def _case(name, seed=11):
    rng = np.random.default_rng(seed)
    if name == "normal":
        return rng.standard_normal((1500, 1500))
    if name == "tall":
        return rng.standard_normal((5000, 700))
    if name == "wide":
        return rng.standard_normal((600, 1400))
    if name == "zero_cols":  # zero columns mid-panel: panel stops, full-pivot fallback, restart
        a = rng.standard_normal((900, 900))
        a[:, [5, 130, 131, 600]] = 0.0
        return a
    if name == "stress":
        return workloads.delay_stress(1024, seed)[0]
    if name == "ties":
        # Exact magnitude ties in many columns (power-of-two multiples of a few values).  Its
        # eliminations cancel to exact zeros or not depending on the order of the sums, which
        # differs between the GPU update and the oracle's BLAS: the two panel kernels are held
        # to each other here, not to the oracle.
        x = rng.integers(-3, 4, size=700).astype(np.float64)
        y = rng.integers(1, 3, size=700).astype(np.float64)
        a = np.outer(2.0 ** x, y)
        a[:, ::3] = np.round(rng.standard_normal((700, 234)))
        return a
    if name == "nan":
        a = rng.standard_normal((640, 640))
        a[rng.integers(0, 640, 9), 3] = np.nan
        a[:, 200] = np.nan
        return a
    if name == "small":
        return rng.standard_normal((37, 23))
    raise KeyError(name)


CASES = ["normal", "tall", "wide", "zero_cols", "stress", "ties", "nan", "small"]


@pytest.mark.parametrize("nb", [64, 128])
@pytest.mark.parametrize("case", CASES)
def test_lag_equals_move_fp64(case, nb):
    a = _case(case)
    lm = _factor(a, dlu.PANEL_MOVE, nb=nb)
    ll = _factor(a, dlu.PANEL_LAG, nb=nb)
    assert ll[3] == lm[3]
    assert np.array_equal(ll[1], lm[1]) and np.array_equal(ll[2], lm[2])
    assert np.array_equal(ll[0], lm[0], equal_nan=True)
    if case not in ("nan", "ties"):
        lu, ipiv, jpiv, info, rep, marg = ll
        check_parity(a, lu, ipiv, jpiv, info, marg)


@pytest.mark.parametrize("case", ["normal", "zero_cols", "ties", "tall"])
def test_lag_equals_move_fp32(case):
    a = _case(case).astype(np.float32).astype(np.float64)
    lm = _factor(a, dlu.PANEL_MOVE, dtype=np.float32)
    ll = _factor(a, dlu.PANEL_LAG, dtype=np.float32)
    assert ll[3] == lm[3]
    assert np.array_equal(ll[1], lm[1]) and np.array_equal(ll[2], lm[2])
    assert np.array_equal(ll[0], lm[0], equal_nan=True)


@pytest.mark.parametrize("rpl", [1, 2, 3, 4])
def test_lag_rows_per_lane(rpl):
    a = _case("zero_cols", seed=5)
    lm = _factor(a, dlu.PANEL_MOVE)
    ll = _factor(a, dlu.PANEL_LAG, rows_per_lane=rpl)
    assert ll[3] == lm[3] and np.array_equal(ll[1], lm[1]) and np.array_equal(ll[2], lm[2])
    assert np.array_equal(ll[0], lm[0])


def test_lag_margins_match_oracle():
    a = _case("normal", seed=3)
    lu, ipiv, jpiv, info, rep, marg = _factor(a, dlu.PANEL_LAG)
    _, _, _, _, ex = oracle.delayed_getrf(a, nb=64, margins=True)
    assert rep["exact_ties"] == 0
    assert abs(rep["min_margin"] - ex["margins"].min()) <= 1e-9
    assert np.allclose(marg, ex["margins"], atol=1e-9)


def test_lag_rejects_what_it_cannot_take():
    import torch
    a = np.random.default_rng(1).standard_normal((300, 300))
    t, lda = dlu.to_device(a)
    st = dlu.getrf_device(300, 300, t.data_ptr(), lda, options={"nb": 256, "panel_kernel": 2})[0]
    assert st == dlu.Status.INVALID_ARGUMENT
    st = dlu.getrf_device(300, 300, t.data_ptr(), lda, options={"panel_kernel": 3})[0]
    assert st == dlu.Status.INVALID_ARGUMENT
    torch.cuda.synchronize()
