"""GPU checks of the FP32 factorization, the solves and the mixed-precision solver (C ABI).

None of these has a counterpart in the reference, which factors only.  What pins them:
  * FP32 factor: exact known answers (small integer matrices, exact in binary32); the zero-column
    invariant and delayed set of the delay-stress input (identical to the FP64 oracle's);
    LAPACK sgetrf (scipy) pivots on inputs without zero columns, up to the first step whose
    reported margin is below the FP32 rounding level; scaled backward error in FP32 units.
  * getrs: factors built by hand with arbitrary row and column swap histories define
    A = P^T (L U) Q^T exactly, so the solve of A x = b is checked against numpy's solve of the
    same A -- this pins the order in which both swap histories are applied; then solves after a
    real factorization, checked by the numpy backward error.
  * backward error kernel: the same quantity computed by numpy.
  * mixed solver: the backward error reaches 1e-13 (checked again by numpy from the returned x)
    and the solution agrees with numpy's FP64 solve to cond * eps.
"""
import numpy as np
import pytest

from helpers import delayed_set, dlu, golden, oracle, workloads

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)
EPS64 = float(np.finfo(np.float64).eps)


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dlu.lib()
    return torch


def dev(a, dtype="float64"):
    """Column-major device copy of a host matrix: (storage tensor of shape (n, m), lda = m)."""
    torch = _torch()
    m, n = a.shape
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dtype).T)).to("cuda")
    return t, max(1, m)


def host(t, m):
    return np.asfortranarray(t.cpu().numpy()[:, :m].T)


def sfactor(a, nb=None):
    torch = _torch()
    m, n = a.shape
    t, lda = dev(a, "float32")
    st, ipiv, jpiv, info, rep, marg = dlu.sgetrf_device(m, n, t.data_ptr(), lda,
                                                         options={"nb": nb or 0}, margins=True)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS, dlu.status_string(st)
    return host(t, m).astype(np.float64), ipiv, jpiv, info, rep, marg


def _case(c):
    a = np.zeros((c["m"], c["n"]), order="F")
    if "dense_rows" in c:
        a[:, :] = np.array(c["dense_rows"])
    for i, j, v in c.get("entries", []):
        a[i, j] = v
    return a


@pytest.mark.parametrize("idx", range(4))
def test_sgetrf_known_answers(idx):
    c = golden("known_answers.json")["cases"][idx]
    a = _case(c)
    lu, ipiv, jpiv, info, rep, _ = sfactor(a)
    assert info == c["info"]
    assert ipiv.tolist() == c["ipiv"] and jpiv.tolist() == c["jpiv"]


def _bwd32(a, lu, ipiv, jpiv, info):
    """||PAQ - LU||_F / (max(m,n) eps32 ||A||_F), evaluated in float64."""
    m, n = a.shape
    return oracle.backward_error(a, lu, ipiv, jpiv, info) * EPS64 / EPS32


@pytest.mark.parametrize("n,seed", [(512, 11), (1000, 12)])
def test_sgetrf_matches_lapack_sgetrf(n, seed):
    import scipy.linalg
    a = workloads.normal(n, n, seed).astype(np.float32).astype(np.float64)
    lu, ipiv, jpiv, info, rep, marg = sfactor(a, nb=128)
    assert info == n
    assert np.array_equal(jpiv, np.arange(n))
    _, piv = scipy.linalg.lu_factor(a.astype(np.float32), check_finite=False)
    diff = np.flatnonzero(ipiv != piv)
    if diff.size:
        k = int(diff[0])
        # rounding in binary32 may legitimately flip a near-tie; it must be reported as one
        assert marg[k] < 1e-3, f"pivots differ at step {k} with margin {marg[k]}"
    assert _bwd32(a, lu, ipiv, jpiv, info) <= 10.0


def test_sgetrf_delay_stress_delayed_set():
    a, zero_cols, _, _ = workloads.delay_stress(512, 3)
    a32 = a.astype(np.float32).astype(np.float64)
    lu, ipiv, jpiv, info, rep, _ = sfactor(a32)
    _, ip_c, jp_c, info_c, _ = oracle.delayed_getrf(a, nb=64)
    assert info == info_c == 512 - len(zero_cols)
    assert np.array_equal(delayed_set(jpiv, info, 512), np.sort(zero_cols))
    assert _bwd32(a32, lu, ipiv, jpiv, info) <= 10.0


def _factors_by_hand(n, seed):
    """Compact L\\U with known L, U and random swap histories; returns (F, ipiv, jpiv, A)."""
    rng = np.random.default_rng(seed)
    low = np.tril(rng.uniform(-0.1, 0.1, (n, n)), -1) + np.eye(n)
    up = np.triu(rng.uniform(-0.1, 0.1, (n, n)), 1) + np.diag(rng.uniform(1.0, 2.0, n))
    f = np.tril(low, -1) + up
    ipiv = np.array([rng.integers(k, n) for k in range(n)], dtype=np.int32)
    jpiv = np.array([rng.integers(k, n) for k in range(n)], dtype=np.int32)
    rows = oracle.replay(ipiv, n, n)
    cols = oracle.replay(jpiv, n, n)
    a = np.zeros((n, n))
    a[np.ix_(rows, cols)] = low @ up  # P A Q = L U
    return np.asfortranarray(f), ipiv, jpiv, a


def _bwd_inf(a, x, b):
    r = b - a @ x
    return (np.abs(r).max(axis=0) /
            (np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0)))


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("n,nrhs", [(300, 3), (129, 17), (64, 1), (1, 1), (65, 40), (3000, 5)])
def test_getrs_applies_both_swap_histories(single, n, nrhs):
    torch = _torch()
    f, ipiv, jpiv, a = _factors_by_hand(n, 100 + n)
    b = np.asfortranarray(np.random.default_rng(n).standard_normal((n, nrhs)))
    dt = "float32" if single else "float64"
    tf, lda = dev(f, dt)
    tb, ldb = dev(b, dt)
    st = dlu.getrs_device(n, nrhs, tf.data_ptr(), lda, ipiv, jpiv, n, tb.data_ptr(), ldb,
                          single=single)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS
    x = host(tb, n).astype(np.float64)
    # Backward error at every size; forward error only while these random unit triangles are
    # well conditioned (their condition number grows with n, to ~1e5 at n = 3000).
    assert _bwd_inf(a, x, b).max() <= (1e-5 if single else 1e-14)
    if n <= 300:
        ref = np.linalg.solve(a, b)
        tol = 1e-4 if single else 1e-12
        assert np.linalg.norm(x - ref) <= tol * np.linalg.norm(ref)


@pytest.mark.parametrize("n,nrhs", [(256, 1), (1000, 3), (2048, 16)])
def test_getrs_after_getrf(n, nrhs):
    torch = _torch()
    a = workloads.normal(n, n, 7 + n)
    b = np.asfortranarray(np.random.default_rng(n).standard_normal((n, nrhs)))
    ta, lda = dev(a)
    st, ipiv, jpiv, info, _, _ = dlu.getrf_device(n, n, ta.data_ptr(), lda)
    assert st == dlu.Status.SUCCESS and info == n
    tb, ldb = dev(b)
    st = dlu.getrs_device(n, nrhs, ta.data_ptr(), lda, ipiv, jpiv, info, tb.data_ptr(), ldb)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS
    x = host(tb, n)
    assert _bwd_inf(a, x, b).max() <= 1e-14


def test_getrs_repeated_calls_are_identical():
    # The triangular solves hand blocks between workgroups through flags tagged with a per-call
    # epoch and a ticket counter that only grows; back-to-back calls of different sizes must
    # not see each other's state.
    torch = _torch()
    outs = []
    for n, nrhs in [(700, 4), (200, 2), (700, 4), (1500, 33), (700, 4)]:
        f, ipiv, jpiv, a = _factors_by_hand(n, 55 + n)
        b = np.asfortranarray(np.random.default_rng(n + nrhs).standard_normal((n, nrhs)))
        tf, lda = dev(f)
        tb, ldb = dev(b)
        st = dlu.getrs_device(n, nrhs, tf.data_ptr(), lda, ipiv, jpiv, n, tb.data_ptr(), ldb)
        torch.cuda.synchronize()
        assert st == dlu.Status.SUCCESS
        x = host(tb, n)
        assert _bwd_inf(a, x, b).max() <= 1e-14
        if n == 700:
            outs.append(x)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_getrs_rejects_a_delayed_factorization():
    torch = _torch()
    a = workloads.normal(200, 200, 5)
    a[:, 17] = 0.0
    ta, lda = dev(a)
    st, ipiv, jpiv, info, _, _ = dlu.getrf_device(200, 200, ta.data_ptr(), lda)
    assert st == dlu.Status.SUCCESS and info == 199
    tb, ldb = dev(np.ones((200, 1)))
    st = dlu.getrs_device(200, 1, ta.data_ptr(), lda, ipiv, jpiv, info, tb.data_ptr(), ldb)
    assert st == dlu.Status.INVALID_ARGUMENT


def test_backward_error_device_matches_numpy():
    _torch()
    n, nrhs = 700, 5
    rng = np.random.default_rng(3)
    a = workloads.normal(n, n, 3)
    x = np.asfortranarray(rng.standard_normal((n, nrhs)))
    # a residual far above the rounding of b - A x, so both summation orders agree closely
    b = np.asfortranarray(a @ x + 1e-6 * rng.standard_normal((n, nrhs)))
    ta, lda = dev(a)
    tx, ldx = dev(x)
    tb, ldb = dev(b)
    st, bwd = dlu.backward_error_device(n, nrhs, ta.data_ptr(), lda, tx.data_ptr(), ldx,
                                        tb.data_ptr(), ldb)
    assert st == dlu.Status.SUCCESS
    ref = _bwd_inf(a, x, b)
    assert np.allclose(bwd, ref, rtol=1e-6, atol=0)


@pytest.mark.parametrize("mixed", [True, False])
def test_gesv_reaches_fp64_backward_error(mixed):
    torch = _torch()
    n, nrhs = 1024, 4
    a = workloads.conditioned(n, 21)
    b = np.asfortranarray(np.random.default_rng(22).standard_normal((n, nrhs)))
    ta, lda = dev(a)
    tb, ldb = dev(b)
    tx = torch.zeros((nrhs, n), dtype=torch.float64, device="cuda")
    st, rep, bwd = dlu.gesv_device(n, nrhs, ta.data_ptr(), lda, tb.data_ptr(), ldb,
                                   tx.data_ptr(), n, mixed=mixed)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS
    assert rep["converged"] == 1 and rep["factor_info"] == n
    assert rep["used_fp64"] == (0 if mixed else 1)
    assert bwd.max() <= 1e-13
    x = host(tx, n)
    assert _bwd_inf(a, x, b).max() <= 1e-13
    ref = np.linalg.solve(a, b)
    # forward error: cond(A) = 1e6, so agreement to about cond * eps
    assert np.linalg.norm(x - ref) <= 1e6 * 1e3 * EPS64 * np.linalg.norm(ref)
    if not mixed:
        assert rep["iterations"] <= 2


def test_gesv_mixed_singular_input_falls_back_and_reports():
    torch = _torch()
    n = 256
    a = workloads.normal(n, n, 9)
    a[:, 40] = 0.0
    ta, lda = dev(a)
    tb, ldb = dev(np.ones((n, 1)))
    tx = torch.zeros((1, n), dtype=torch.float64, device="cuda")
    st, rep, _ = dlu.gesv_device(n, 1, ta.data_ptr(), lda, tb.data_ptr(), ldb, tx.data_ptr(), n,
                                 mixed=True)
    assert st == dlu.Status.INVALID_ARGUMENT
    assert rep["factor_info"] == n - 1 and rep["used_fp64"] == 1


@pytest.mark.parametrize("n,seed", [(6000, 21), (8000, 22)])
def test_sgetrf_default_path_above_4096(n, seed):
    """The FP32 defaults above 4096 rows: a CU-partitioned panel stream, the far GEMM's tail on
    the panel stream, 128 x 128 FP32 GEMM tiles and the pair schedule.  Pivots equal LAPACK
    sgetrf's up to a reported FP32 near-tie; the factor's backward error is <= 10 in FP32 units."""
    a = workloads.normal(n, n, seed).astype(np.float32).astype(np.float64)
    lu, ipiv, jpiv, info, rep, marg = sfactor(a)
    assert info == n
    assert np.array_equal(jpiv, np.arange(n))
    _, piv, infol = oracle.lapack_sgetrf(a)
    diff = np.flatnonzero(ipiv != piv)
    if diff.size:
        k = int(diff[0])
        print(f"n={n}: first difference from LAPACK sgetrf at step {k}, margin {marg[k]:.3e}")
        assert marg[k] < 1e-3, f"pivots differ at step {k} with margin {marg[k]}"
    b32 = _bwd32(a, lu, ipiv, jpiv, info)
    print(f"n={n}: FP32 factor backward error {b32:.3f} (FP32 units)", flush=True)
    assert b32 <= 10.0
