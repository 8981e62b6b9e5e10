"""The allocation-free entry point (dlu_getrf_device_noalloc) against dlu_getrf_device, through
the C ABI: two calls in a row, the second smaller, each giving the same info, ipiv, jpiv and
delayed set bit for bit and factors within 1e-12 relative (Frobenius; the two entry points block
the trailing update differently); sizes beyond the scratch array are refused on the host."""
import ctypes

import numpy as np
import pytest

from helpers import delayed_set, dlu, oracle, workloads

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dlu.lib()
    return torch


def _call(fn, a):
    torch = _torch()
    m, n = a.shape
    t, lda = dlu.to_device(a)
    p = min(m, n)
    ipiv = np.zeros(p, dtype=np.int32)
    jpiv = np.zeros(p, dtype=np.int32)
    info = ctypes.c_int(-7)
    st = fn(m, n, t.data_ptr(), lda, ipiv.ctypes.data_as(dlu._ip), jpiv.ctypes.data_as(dlu._ip),
            ctypes.byref(info), None)
    torch.cuda.synchronize()
    return dlu.Status(st), dlu.from_device(t, m), ipiv, jpiv, info.value


def test_noalloc_matches_getrf_device():
    L = dlu.lib()
    big = workloads.normal(1500, 1500, 71)
    big[:, [3, 400, 401, 1200]] = 0.0  # fallbacks, restarts and a delayed suffix
    small = workloads.delay_stress(700, 72)[0]
    for a in (big, small):
        m, n = a.shape
        print(f"noalloc call m={m} n={n}", flush=True)
        st0, lu0, ip0, jp0, info0 = _call(L.dlu_getrf_device, a)
        st1, lu1, ip1, jp1, info1 = _call(L.dlu_getrf_device_noalloc, a)
        print(f"statuses: getrf_device {st0!r}, noalloc {st1!r}", flush=True)
        assert st0 == dlu.Status.SUCCESS and st1 == dlu.Status.SUCCESS
        assert info1 == info0
        assert np.array_equal(ip1, ip0) and np.array_equal(jp1, jp0)
        assert np.array_equal(delayed_set(jp1, info1, n), delayed_set(jp0, info0, n))
        assert np.linalg.norm(lu1 - lu0) <= 1e-12 * np.linalg.norm(lu0)
        assert np.all(lu1[info1:, info1:] == lu0[info0:, info0:])
        assert oracle.backward_error(a, lu1, ip1, jp1, info1) <= 10.0


def test_noalloc_rejects_sizes_beyond_its_scratch():
    L = dlu.lib()
    ip = np.zeros(1, dtype=np.int32)
    info = ctypes.c_int(7)
    st = L.dlu_getrf_device_noalloc(40000, 40000, 1234, 40000, ip.ctypes.data_as(dlu._ip),
                                    ip.ctypes.data_as(dlu._ip), ctypes.byref(info), None)
    assert st == dlu.Status.INVALID_ARGUMENT and info.value == -1
