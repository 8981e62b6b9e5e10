"""Left row interchanges deferred to one composed pass at the end, against interchanging them as
each panel ends (dlu_options.defer_left < 0).

The deferral moves the same values to the same rows, only later, so the two factorizations must
agree bit for bit.  The cases reach every place a panel's movement is recorded: single panels,
panel pairs (n above the pairing threshold), rectangular shapes, a full-pivot fallback mid-run
(the owed movements are applied before the fallback interchanges whole rows) and the FP32 factor.
"""
import numpy as np
import pytest

from helpers import dlu, workloads
from test_gpu_getrf import check_parity

pytestmark = pytest.mark.gpu


def _factor(a, defer, dtype=np.float64, **opts):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    m, n = a.shape
    t, lda = dlu.to_device(a)
    options = dict(opts, defer_left=0 if defer else -1)
    if dtype == np.float32:
        t = t.to(torch.float32)
        out = dlu.sgetrf_device(m, n, t.data_ptr(), lda, options=options, margins=True)
    else:
        out = dlu.getrf_device(m, n, t.data_ptr(), lda, options=options, margins=True)
    torch.cuda.synchronize()
    st, ipiv, jpiv, info, rep, marg = out
    assert st == dlu.Status.SUCCESS, dlu.status_string(st)
    return np.asfortranarray(t.double().cpu().numpy()[:, :m].T), ipiv, jpiv, info, marg


def _same(x, y):
    assert x[3] == y[3]
    assert np.array_equal(x[1], y[1]) and np.array_equal(x[2], y[2])
    assert np.array_equal(x[0], y[0], equal_nan=True)


@pytest.mark.parametrize("shape,nb", [((2000, 2000), 128), ((4500, 4500), 128), ((3000, 3000), 64),
                                      ((6000, 1000), 128), ((900, 3000), 128), ((257, 255), 64)])
def test_deferred_equals_immediate(shape, nb):
    a = workloads.normal(shape[0], shape[1], 7 + shape[0])
    on, off = _factor(a, True, nb=nb), _factor(a, False, nb=nb)
    _same(on, off)
    check_parity(a, on[0], on[1], on[2], on[3], on[4])


def test_deferred_with_fallback_mid_run():
    # Zero columns past the first panels: the deferred movements are applied at the first
    # fallback, the rest of the run interchanges immediately.
    a = workloads.normal(1500, 1500, 11)
    a[:, [700, 701, 1200]] = 0.0
    on, off = _factor(a, True), _factor(a, False)
    _same(on, off)
    check_parity(a, on[0], on[1], on[2], on[3], on[4])


def test_deferred_delay_stress():
    a = workloads.delay_stress(2048, 3)[0]
    on, off = _factor(a, True), _factor(a, False)
    _same(on, off)


def test_deferred_fp32():
    a = workloads.normal(3500, 3500, 5).astype(np.float32).astype(np.float64)
    _same(_factor(a, True, dtype=np.float32), _factor(a, False, dtype=np.float32))


def test_deferred_at_its_row_limit_and_beyond():
    # 16384 rows is the largest deferred case; taller matrices interchange immediately.
    a = workloads.normal(16384, 512, 13)
    _same(_factor(a, True), _factor(a, False))
    b = workloads.normal(16390, 300, 14)
    _same(_factor(b, True), _factor(b, False))
