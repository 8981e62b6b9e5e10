"""GPU parity of the delayed Bunch-Kaufman LDL^T (delayed-bunch-kaufman/delayed_sytrf_blocked.cpp:
168-358) against the oracle's restatement of it, through the C ABI.

  * the reference's own 4096-node example (known answer derived by hand, as in test_oracle);
  * random symmetric matrices with and without zero nodes, at the reference's panel width 64 and
    at 16 and 2 (panel boundaries, the 2x2 break): perm, piv_size and info bit-exact against the
    blocked oracle, the factor within 1e-12 relative (Frobenius, lower triangle), and
    P^T A P = L D L^T to 1e-13 relative;
  * the upper triangle is neither read nor written; the host interface equals the device one.
A pivot decision is accepted as different only at a Bunch-Kaufman test whose two sides are
within 1e-10 relative of each other (a near-tie that rounding may flip); none is expected on
these inputs.
"""
import numpy as np
import pytest

from helpers import dlu, oracle

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dlu.lib()
    return torch


def _sym(n, seed, zero_nodes=(), scale_nodes=()):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n))
    s = np.asfortranarray(x + x.T)
    z = list(zero_nodes)
    s[z, :] = 0.0
    s[:, z] = 0.0
    for i in scale_nodes:  # tiny diagonals: the 2x2 and interchange tests fire more often
        s[i, i] *= 1e-3
    return s


def gpu_sytrf(a, nb=64):
    torch = _torch()
    n = a.shape[0]
    t, lda = dlu.to_device(a)
    st, perm, piv, info = dlu.sytrf_device(n, t.data_ptr(), lda, nb=nb)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS, dlu.status_string(st)
    return dlu.from_device(t, n), perm, piv, info


def _check(a, nb=64):
    f, perm, piv, info = gpu_sytrf(a, nb)
    fc, pc, qc, ic = oracle.delayed_sytrf(a, nb=nb)
    assert info == ic
    assert np.array_equal(perm, pc) and np.array_equal(piv, qc), np.flatnonzero(perm != pc)[:5]
    lo, loc = np.tril(f), np.tril(fc)
    assert np.linalg.norm(lo - loc) <= 1e-12 * max(np.linalg.norm(loc), 1e-300)
    assert oracle.sytrf_reconstruction_error(a, f, perm, piv, info) <= 1e-13
    return f, perm, piv, info


def test_sytrf_known_answer_example_cpp():
    n = 4096
    a = np.zeros((n, n), order="F")
    a[3, 2] = a[2, 3] = 2.0
    for k in range(6):
        a[k + 4, k + 4] = 6.0 - k
    f, perm, piv, info = gpu_sytrf(a)
    assert info == 8
    assert perm.tolist() == [9, 8, 2, 3, 4, 5, 6, 7, 1] + list(range(10, n)) + [0]
    assert piv.tolist() == [1, 1, 2, 0, 1, 1, 1, 1] + [0] * (n - 8)
    assert [f[0, 0], f[1, 1], f[2, 2], f[3, 2], f[3, 3]] == [1.0, 2.0, 0.0, 2.0, 0.0]
    assert [f[k, k] for k in range(4, 8)] == [6.0, 5.0, 4.0, 3.0]
    assert not np.tril(f, -1)[4:, :].any() and not f[8:, 8:].any()


@pytest.mark.parametrize("n,seed,nb", [(50, 1, 64), (200, 2, 64), (700, 3, 64), (1500, 4, 64),
                                       (333, 5, 16), (130, 6, 2), (2048, 7, 64)])
def test_sytrf_matches_oracle(n, seed, nb):
    a = _sym(n, seed)
    _, _, piv, info = _check(a, nb)
    assert info == n and (piv == 2).sum() > 0


@pytest.mark.parametrize("n,seed,nz,nb", [(120, 11, 12, 64), (500, 12, 50, 64), (900, 13, 90, 16),
                                          (64, 14, 64, 64), (300, 15, 1, 64)])
def test_sytrf_zero_nodes_are_delayed(n, seed, nz, nb):
    z = np.random.default_rng(seed).choice(n, size=nz, replace=False)
    a = _sym(n, seed, z, scale_nodes=range(0, n, 7))
    f, perm, piv, info = _check(a, nb)
    assert info == n - nz
    assert sorted(perm[info:].tolist()) == sorted(z.tolist())


def test_sytrf_upper_triangle_untouched_and_host_equals_device():
    n = 260
    a = _sym(n, 21)
    iu = np.triu_indices(n, 1)
    a[iu] = -3.5  # not the symmetric values
    f, perm, piv, info = gpu_sytrf(a)
    assert np.all(f[iu] == -3.5)
    h = np.asfortranarray(a.copy())
    import ctypes
    ph = np.zeros(n, dtype=np.int32)
    qh = np.zeros(n, dtype=np.int32)
    ih = ctypes.c_int(-7)
    st = dlu.lib().dlu_sytrf_host(n, h.ctypes.data_as(ctypes.c_void_p), n,
                                  ph.ctypes.data_as(dlu._ip), qh.ctypes.data_as(dlu._ip),
                                  ctypes.byref(ih), None)
    assert st == dlu.Status.SUCCESS and ih.value == info
    assert np.array_equal(ph, perm) and np.array_equal(qh, piv) and np.array_equal(h, f)


def test_sytrf_rejects_bad_arguments():
    import ctypes
    L = dlu.lib()
    p = np.zeros(4, dtype=np.int32)
    info = ctypes.c_int(7)
    assert L.dlu_sytrf_device(4, None, 4, p.ctypes.data_as(dlu._ip), p.ctypes.data_as(dlu._ip),
                              ctypes.byref(info), None) == dlu.Status.INVALID_ARGUMENT
    assert info.value == -1
    assert L.dlu_sytrf_device(4, 1234, 3, p.ctypes.data_as(dlu._ip), p.ctypes.data_as(dlu._ip),
                              ctypes.byref(info), None) == dlu.Status.INVALID_ARGUMENT
    assert L.dlu_sytrf_device(0, None, 1, None, None, ctypes.byref(info), None) == dlu.Status.SUCCESS
    assert info.value == 0
