"""GPU parity of the column-delay variant and of tall matrices, through the C ABI.

  * the column-delay variant (delayed_lu/delayed_dgetrf_column.cpp:99-211) against the oracle's
    restatement of it: the reference's own rank-800 example (known answer, derived by hand),
    N(0,1) matrices with zero columns, the delay-stress recipe, rectangular shapes, exact ties,
    NaN columns;
  * matrices with more rows than the register-resident panel kernel holds (m > 65536 on 256 CUs):
    the tall path must give the oracle's pivots bit for bit (delayed_dgetrf_cuda.cu:385-397 puts
    no limit on m), with and without zero columns, in both variants.

Bars as in test_gpu_getrf: ipiv / jpiv / info / delayed set bit-exact (a difference is allowed
only at a step whose reported margin is a near-tie), factors within 1e-12 relative (Frobenius),
scaled backward error <= 10.
"""
import numpy as np
import pytest

from helpers import delayed_set, dlu, oracle, workloads

pytestmark = pytest.mark.gpu

FACTOR_RTOL = 1e-12


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dlu.lib()
    return torch


def gpu_factor(a, column, nb=None, sched=None):
    torch = _torch()
    m, n = a.shape
    t, lda = dlu.to_device(a)
    opts = dict(sched or {})
    opts["fallback"] = dlu.FALLBACK_COLUMN if column else dlu.FALLBACK_FULL
    if nb:
        opts["nb"] = nb
    st, ipiv, jpiv, info, rep, marg = dlu.getrf_device(m, n, t.data_ptr(), lda, options=opts,
                                                        margins=True)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS, dlu.status_string(st)
    lu = dlu.from_device(t, m)
    del t
    return lu, ipiv, jpiv, info, rep, marg


def check_parity(a, lu, ipiv, jpiv, info, marg, column):
    lu_c, ip_c, jp_c, info_c, ex = oracle.delayed_getrf(a, nb=64, margins=True, column=column)
    same = info == info_c and np.array_equal(ipiv, ip_c) and np.array_equal(jpiv, jp_c)
    if not same:
        k = int(np.argmax((ipiv != ip_c) | (jpiv != jp_c)))
        pytest.fail(f"pivots differ first at step {k}: gpu ({ipiv[k]},{jpiv[k]}) cpu "
                    f"({ip_c[k]},{jp_c[k]}); margins gpu={marg[k]} cpu={ex['margins'][k]}; "
                    f"info {info} vs {info_c}")
    assert np.linalg.norm(lu - lu_c) <= FACTOR_RTOL * np.linalg.norm(lu_c)
    bwd = oracle.backward_error(a, lu, ipiv, jpiv, info)
    assert 0 <= bwd <= 10.0
    return ex


def test_column_known_answer_example_cpp():
    """delayed_lu/example.cpp:204-253 run through the column variant (device and host entry
    points): steps 0 and 1 take the rightmost nonzero columns 801 and 800, info = 800."""
    a = workloads.diag_rank(4096, 800)
    want = [801, 800] + list(range(2, 800)) + list(range(800, 4096))
    lu, ipiv, jpiv, info, rep, _ = gpu_factor(a, column=True)
    assert info == 800 and ipiv.tolist() == want and jpiv.tolist() == want
    assert rep["fallbacks"] == 2  # applied fallbacks; the oracle also counts the final search
    assert np.all(lu[800:, 800:] == 0.0)
    st, lu_h, ip_h, jp_h, info_h = dlu.getrf_host(a, variant="column")
    assert st == dlu.Status.SUCCESS and info_h == 800
    assert ip_h.tolist() == want and jp_h.tolist() == want
    assert np.array_equal(lu_h, lu)


@pytest.mark.parametrize("cols", [[0], [5, 6], [127, 128, 129], [300, 10, 700], [767]])
def test_column_zero_columns(cols):
    n = 768
    a = workloads.normal(n, n, 500 + sum(cols))
    a[:, cols] = 0.0
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, column=True, nb=128)
    assert info == n - len(cols)
    assert np.array_equal(delayed_set(jpiv, info, n), np.sort(cols))
    check_parity(a, lu, ipiv, jpiv, info, marg, column=True)


@pytest.mark.parametrize("n,seed,nb", [(512, 3, 128), (1024, 5, 64), (2048, 11, 128)])
def test_column_delay_stress(n, seed, nb):
    a, zc, _, _ = workloads.delay_stress(n, seed)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, column=True, nb=nb)
    assert info == n - len(zc)
    assert np.array_equal(delayed_set(jpiv, info, n), zc)
    assert np.all(lu[info:, info:] == 0.0)
    check_parity(a, lu, ipiv, jpiv, info, marg, column=True)


@pytest.mark.parametrize("m,n", [(900, 500), (500, 900)])
def test_column_rectangular(m, n):
    a = workloads.normal(m, n, m + 7 * n)
    a[:, [0, 3, 200, 499]] = 0.0
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, column=True, nb=128)
    check_parity(a, lu, ipiv, jpiv, info, marg, column=True)


def test_column_exact_ties_and_nan():
    """Column 0 zero: the rightmost column (n-1) holds two entries of equal magnitude; the
    lower row wins (idamax).  A column of NaN and zeros right of it counts as zero (NaN never
    compares greater), so it is skipped by the search and ends up delayed."""
    n = 600
    a = workloads.normal(n, n, 61)
    a[:, 0] = 0.0
    a[:, n - 1] = 0.0
    a[100, n - 1] = -4.0
    a[450, n - 1] = 4.0
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, column=True)
    assert (int(ipiv[0]), int(jpiv[0])) == (100, n - 1)
    check_parity(a, lu, ipiv, jpiv, info, marg, column=True)
    b = workloads.normal(n, n, 62)
    b[:, 0] = 0.0
    b[:, n - 1] = 0.0
    b[7, n - 1] = np.nan
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(b, column=True)
    assert int(jpiv[0]) == n - 2
    assert info < n and (n - 1) in delayed_set(jpiv, info, n).tolist()
    assert np.isfinite(lu[:info, :info]).all()


@pytest.mark.parametrize("m,n,zero,column", [(100000, 128, [], False), (70000, 300, [], False),
                                             (70000, 300, [5, 130, 131, 299], False),
                                             (70000, 300, [5, 130, 131, 299], True)])
def test_tall_matrices(m, n, zero, column):
    """More rows than the register-resident panel kernel holds: the tall path (one column step
    per launch sequence, whole-row interchanges) gives the oracle's pivots bit for bit."""
    a = workloads.normal(m, n, m + n + len(zero))
    a[:, zero] = 0.0
    print(f"tall m={m} n={n} zero={zero} column={column}", flush=True)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, column=column)
    assert info == n - len(zero)
    if zero:
        assert np.array_equal(delayed_set(jpiv, info, n), np.sort(zero))
    check_parity(a, lu, ipiv, jpiv, info, marg, column=column)


def test_tall_then_regular_panels():
    """A square matrix taller than the panel kernel's capacity: the first panels go through the
    tall path, the later ones (fewer rows left) through the register-resident kernel, with
    look-ahead resuming once the next panel fits."""
    m, n = 66000, 640
    a = workloads.normal(m, n, 66)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, column=False, nb=128)
    assert info == n
    check_parity(a, lu, ipiv, jpiv, info, marg, column=False)
