"""GPU parity of the HIP delayed LU against the oracle, through the C ABI.

Bars: ipiv / jpiv / info / delayed set bit-exact versus the CPU oracle
(nb = 64, the reference's blocking) unless a near-tie is reported at the first differing step;
factors within 1e-12 relative in the Frobenius norm, ||F_gpu - F_cpu||_F / ||F_cpu||_F (the GPU
blocks by 256 and sums in MFMA order, the oracle by 64 in BLAS order, so rounding differs at the
unit-roundoff level, amplified by element growth); scaled backward error
||PAQ - LU||_F / (max(m,n) eps ||A||_F) <= 10.
"""
import sys

import numpy as np
import pytest

from helpers import delayed_set, dlu, golden, oracle, reconstruction_error, workloads

pytestmark = pytest.mark.gpu

FACTOR_RTOL = 1e-12
BWD_MAX = 10.0


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dlu.lib()  # raises when libdlu.so is missing
    return torch


def gpu_factor(a, nb=None, lda=None, lookahead=None, margins=True, sched=None):
    torch = _torch()
    m, n = a.shape
    t, lda = dlu.to_device(a, lda)
    opts = dict(sched or {})
    if nb:
        opts["nb"] = nb
    if lookahead is not None:
        opts["lookahead"] = lookahead
    st, ipiv, jpiv, info, rep, marg = dlu.getrf_device(m, n, t.data_ptr(), lda, options=opts,
                                                        margins=margins)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS, dlu.status_string(st)
    return dlu.from_device(t, m), ipiv, jpiv, info, rep, marg


def check_parity(a, lu, ipiv, jpiv, info, marg, nb_cpu=64):
    lu_c, ip_c, jp_c, info_c, ex = oracle.delayed_getrf(a, nb=nb_cpu, margins=True)
    same = info == info_c and np.array_equal(ipiv, ip_c) and np.array_equal(jpiv, jp_c)
    if not same:
        k = int(np.argmax((ipiv != ip_c) | (jpiv != jp_c)))
        gm = marg[k] if marg is not None else None
        cm = ex["margins"][k]
        pytest.fail(f"pivots differ first at step {k}: gpu ({ipiv[k]},{jpiv[k]}) cpu "
                    f"({ip_c[k]},{jp_c[k]}); margins gpu={gm} cpu={cm}; info {info} vs {info_c}")
    assert np.linalg.norm(lu - lu_c) <= FACTOR_RTOL * np.linalg.norm(lu_c)
    bwd = oracle.backward_error(a, lu, ipiv, jpiv, info)
    assert 0 <= bwd <= BWD_MAX
    return bwd


def _case_matrix(c):
    a = np.zeros((c["m"], c["n"]), order="F")
    if "dense_rows" in c:
        a[:, :] = np.array(c["dense_rows"])
    for i, j, v in c.get("entries", []):
        a[i, j] = v
    return a


@pytest.mark.parametrize("config", [-1, 0, 1, 2, 3])
@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (64, 64, 16), (100, 37, 13), (513, 260, 256),
                                   (1024, 1024, 256), (2000, 3000, 64), (128, 128, 1)])
def test_gemm_sub_matches_numpy(M, N, K, config):
    torch = _torch()
    rng = np.random.default_rng(M * 7 + N + K)
    ld = M + 3
    c = rng.standard_normal((M, N))
    a = rng.standard_normal((M, K))
    b = rng.standard_normal((K, N))
    tc, _ = dlu.to_device(c, ld)
    ta, _ = dlu.to_device(a, ld)
    tb, _ = dlu.to_device(b, ld)
    st = dlu.gemm_sub_device(M, N, K, tc.data_ptr(), ta.data_ptr(), tb.data_ptr(), ld,
                             config=config)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS
    got = dlu.from_device(tc, M)
    ref = c - a @ b
    assert np.abs(got - ref).max() <= 1e-13 * K * max(1.0, np.abs(ref).max())
    # padding rows between M and ld untouched
    assert np.all(tc.cpu().numpy()[:, M:] == 0.0)


@pytest.mark.parametrize("idx", range(4))
@pytest.mark.parametrize("variant", ["host", "native", "hybrid"])
def test_known_answers_host_interface(idx, variant):
    _torch()
    c = golden("known_answers.json")["cases"][idx]
    a = _case_matrix(c)
    st, lu, ipiv, jpiv, info = dlu.getrf_host(a, variant=variant)
    assert st == dlu.Status.SUCCESS
    assert info == c["info"]
    assert ipiv.tolist() == c["ipiv"] and jpiv.tolist() == c["jpiv"]
    tol = 1e-11 * max(1.0, np.abs(a).max()) * max(a.shape)  # example.cu:101-106
    assert reconstruction_error(a, lu, ipiv, jpiv, info) <= tol


def test_known_answer_example_cpp_device():
    c = golden("known_answers.json")["cases"][4]
    a = _case_matrix(c)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a)
    assert info == 800
    assert ipiv.tolist() == c["ipiv"] and jpiv.tolist() == c["jpiv"]
    assert rep["fallbacks"] == 800
    # the delayed Schur complement is the untouched zero block
    assert np.all(lu[800:, 800:] == 0.0)


@pytest.mark.parametrize("n,nb", [(1, 256), (2, 256), (17, 64), (64, 64), (100, 128),
                                  (257, 256), (512, 64), (512, 256), (1000, 128), (2048, 256)])
def test_normal_parity(n, nb):
    a = workloads.normal(n, n, 1000 + n)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, nb=nb)
    check_parity(a, lu, ipiv, jpiv, info, marg)
    assert rep["fallbacks"] == 0


@pytest.mark.parametrize("m,n", [(300, 200), (200, 300), (1000, 64), (64, 1000), (777, 513)])
def test_rectangular_parity(m, n):
    a = workloads.normal(m, n, m * 31 + n)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, nb=128)
    check_parity(a, lu, ipiv, jpiv, info, marg)


def test_leading_dimension_padding():
    a = workloads.normal(300, 280, 9)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, lda=333)
    check_parity(a, lu, ipiv, jpiv, info, marg)


@pytest.mark.parametrize("n,seed,nb", [(512, 3, 256), (1024, 5, 256), (1024, 5, 64),
                                       (2048, 11, 256)])
def test_delay_stress_parity(n, seed, nb):
    a, zc, tc, off = workloads.delay_stress(n, seed)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, nb=nb)
    assert info == n - len(zc)
    assert np.array_equal(delayed_set(jpiv, info, n), zc)
    assert np.all(lu[info:, info:] == 0.0)
    check_parity(a, lu, ipiv, jpiv, info, marg)


@pytest.mark.parametrize("cols", [[0], [5], [100, 101], [255, 256, 257], [300, 10, 700]])
def test_zero_columns_mid_panel(cols):
    n = 768
    a = workloads.normal(n, n, 77 + sum(cols))
    a[:, cols] = 0.0
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, nb=256)
    assert info == n - len(cols)
    assert np.array_equal(delayed_set(jpiv, info, n), np.sort(cols))
    check_parity(a, lu, ipiv, jpiv, info, marg)


def test_oracle_golden_vectors_on_gpu():
    d = golden("oracle_vectors.json")
    for c in d["cases"]:
        if c["kind"] == "normal":
            a = workloads.normal(c["m"], c["n"], c["seed"])
        else:
            a = workloads.delay_stress(c["n"], c["seed"])[0]
        print("case", c["name"], flush=True)
        lu, ipiv, jpiv, info, rep, marg = gpu_factor(a)
        assert info == c["info"], c["name"]
        assert ipiv.tolist() == c["ipiv"], c["name"]
        assert jpiv.tolist() == c["jpiv"], c["name"]


def test_all_zero_and_rank_one():
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(np.zeros((300, 300), order="F"))
    assert info == 0 and ipiv.tolist() == list(range(300))
    # powers of two keep every product, reciprocal and difference exact: the Schur complement
    # after one step is exactly zero, so the rest is delayed
    x = 2.0 ** (np.arange(300) % 10)
    y = 2.0 ** (np.arange(300) % 7 - 3)
    a = np.asfortranarray(np.outer(x, y))
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a)
    check_parity(a, lu, ipiv, jpiv, info, marg)
    assert info == 1


def test_nan_column_is_never_a_pivot():
    """GPU rule (delayed_lu_cuda/delayed_dgetrf_cuda.cu:46): NaN never wins a comparison, so an
    all-NaN column is treated like a zero column and ends up delayed.  The CPU reference depends
    on its BLAS idamax here, so this is checked on the GPU alone."""
    n = 200
    a = workloads.normal(n, n, 5)
    a[:, 37] = np.nan
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a)
    assert info == n - 1
    assert delayed_set(jpiv, info, n).tolist() == [37]
    assert np.isfinite(lu[:info, :info]).all()


def test_full_pivot_search_exact_ties():
    """Full-pivot fallback (delayed_dgetrf.cpp:32-47): among equal largest magnitudes the lowest
    column-major offset of the active block wins.  Column 0 is zero, so step 0 searches the whole
    1500 x 1500 block; three entries of magnitude 7.5 (above every N(0,1) entry of this matrix)
    sit in different columns and 512-row segments, i.e. in different waves of the search."""
    n = 1500
    a = workloads.normal(n, n, 11)
    assert np.abs(a).max() < 7.5
    a[:, 0] = 0.0
    a[1200, 9] = -7.5
    a[1499, 3] = 7.5
    a[700, 3] = -7.5
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a)
    assert (int(ipiv[0]), int(jpiv[0])) == (700, 3)
    assert rep["fallbacks"] >= 1
    check_parity(a, lu, ipiv, jpiv, info, marg)


def test_full_pivot_search_skips_nan():
    """A NaN in the active block never wins the full-pivot search (NaN compares false,
    delayed_lu_cuda/delayed_dgetrf_cuda.cu:46): the first fallback takes the largest finite
    entry although the NaN lies at a lower offset."""
    n = 1500
    a = workloads.normal(n, n, 12)
    assert np.abs(a).max() < 7.5
    a[:, 0] = 0.0
    a[3, 1] = np.nan
    a[900, 6] = 7.5
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a)
    assert (int(ipiv[0]), int(jpiv[0])) == (900, 6)


def test_concurrent_factorizations_match_alone():
    """Factorizations called from several host threads at once, each on its own stream, overlap
    on the device (each call leases its own workspace; panels are admitted while their
    workgroups fit on the CUs).  Every result equals the same factorization run alone, bit for
    bit, and the calls raise no error."""
    import threading
    torch = _torch()
    mats = [workloads.normal(n, n, seed) for n, seed in ((1024, 21), (1500, 22), (2048, 23),
                                                        (700, 24), (1800, 25))]
    alone = [gpu_factor(a, margins=False) for a in mats]
    dev = [dlu.to_device(a) for a in mats]
    streams = [torch.cuda.Stream() for _ in mats]
    torch.cuda.synchronize()
    out = [None] * len(mats)

    def work(i):
        t, lda = dev[i]
        m, n = mats[i].shape
        out[i] = dlu.getrf_device(m, n, t.data_ptr(), lda, stream=streams[i])

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(mats))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    for i, a in enumerate(mats):
        st, ipiv, jpiv, info, _, _ = out[i]
        assert st == dlu.Status.SUCCESS, dlu.status_string(st)
        lu = dlu.from_device(dev[i][0], a.shape[0])
        lu0, ipiv0, jpiv0, info0 = alone[i][:4]
        assert info == info0
        assert np.array_equal(ipiv, ipiv0) and np.array_equal(jpiv, jpiv0)
        assert np.array_equal(lu, lu0)


def test_host_equals_device_and_deterministic():
    a = workloads.normal(700, 700, 21)
    lu1, ip1, jp1, info1, _, _ = gpu_factor(a, margins=False)
    lu2, ip2, jp2, info2, _, _ = gpu_factor(a, margins=False)
    st, lu3, ip3, jp3, info3 = dlu.getrf_host(a)
    assert st == dlu.Status.SUCCESS
    assert np.array_equal(lu1, lu2) and np.array_equal(lu1, lu3)
    assert np.array_equal(ip1, ip3) and np.array_equal(jp1, jp3) and info1 == info3


def test_lookahead_off_gives_same_result():
    a = workloads.normal(1500, 1500, 33)
    on = gpu_factor(a, nb=256)
    off = gpu_factor(a, nb=256, lookahead=-1)
    assert np.array_equal(on[1], off[1]) and np.array_equal(on[2], off[2])
    assert np.abs(on[0] - off[0]).max() <= 1e-13 * np.abs(on[0]).max()


def test_margins_report_matches_oracle():
    a = workloads.normal(512, 512, 1)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a)
    _, _, _, _, ex = oracle.delayed_getrf(a, margins=True)
    assert rep["exact_ties"] == 0
    assert abs(rep["min_margin"] - ex["margins"].min()) <= 1e-9
    assert np.allclose(marg, ex["margins"], atol=1e-9)


@pytest.mark.parametrize("n", [4096, 8192])
def test_large_normal_parity(n):
    a = workloads.normal(n, n, 20260000 + n)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, nb=256)
    print(f"n={n} panels={rep['panels']} min_margin={rep['min_margin']:.3e}", flush=True)
    lu_c, ip_c, jp_c, info_c, ex = oracle.delayed_getrf(a, nb=64)
    assert info == info_c == n
    if not np.array_equal(ipiv, ip_c):
        k = int(np.argmax(ipiv != ip_c))
        assert marg[k] < 1e-10, f"step {k} differs with margin {marg[k]}"
    assert np.array_equal(jpiv, jp_c)
    assert np.linalg.norm(lu - lu_c) <= FACTOR_RTOL * np.linalg.norm(lu_c)


# Update schedules: panel pairs sharing a K = 2 nb far update on a third stream (merge_rows=1
# pairs every eligible panel), and the far GEMM's last columns moved to the panel stream
# (split_update > 0 forces it at these sizes).  Every schedule must give the oracle's result.
SCHEDULES = [{"merge_rows": 1}, {"merge_rows": 1, "split_update": 384}, {"split_update": 384},
             {"merge_rows": 1, "split_update": -1}]


@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda d: ",".join(f"{k}={v}" for k, v in d.items()))
@pytest.mark.parametrize("case", ["normal", "rect_tall", "rect_wide", "stress", "zero_cols"])
def test_schedules_match_oracle(sched, case):
    if case == "normal":
        a = workloads.normal(1300, 1300, 404)
    elif case == "rect_tall":
        a = workloads.normal(1500, 900, 405)
    elif case == "rect_wide":
        a = workloads.normal(900, 1400, 406)
    elif case == "stress":
        a, zc, _, _ = workloads.delay_stress(1024, 7)
    else:
        a = workloads.normal(1200, 1200, 407)
        a[:, [300, 301, 650, 1000]] = 0.0  # breaks pairs mid-way
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, nb=128, sched=sched)
    check_parity(a, lu, ipiv, jpiv, info, marg)
    ref = gpu_factor(a, nb=128)
    assert np.array_equal(ipiv, ref[1]) and np.array_equal(jpiv, ref[2]) and info == ref[3]


def test_stream_pairs_of_other_partitions_are_replaced():
    """Four FP32 factorizations with four different panel CU reservations fill a workspace's four
    stream pairs with four partitions.  An FP64 call on the same thread asks for no partition: it
    must get an unpartitioned pair (a replaced slot), not a partitioned one whose panel stream is
    too small for its panel workgroups."""
    torch = _torch()
    for n, cus in ((5000, 32), (10000, 64), (16384, 96), (20000, 128)):
        t = torch.randn((n, n), dtype=torch.float32, device="cuda")
        st, ipiv, jpiv, info, rep, _ = dlu.sgetrf_device(n, n, t.data_ptr(), n,
                                                         options={"panel_cus": cus})
        torch.cuda.synchronize()
        assert st == dlu.Status.SUCCESS and info == n, (n, dlu.status_string(st))
        del t
    n = 16384
    t = torch.randn((n, n), dtype=torch.float64, device="cuda")
    st, ipiv, jpiv, info, rep, _ = dlu.getrf_device(n, n, t.data_ptr(), n)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS and info == n, dlu.status_string(st)
    del t
    torch.cuda.empty_cache()


@pytest.mark.parametrize("panel_cus", [200, 240])
def test_fallback_restart_with_forced_panel_cus(panel_cus):
    """With the panel stream given most of the CUs, the update stream keeps too few for a panel's
    workgroups: the panel restarted after a full-pivot fallback must run on the panel stream."""
    a, zc, _, _ = workloads.delay_stress(4096, 17)
    lu, ipiv, jpiv, info, rep, marg = gpu_factor(a, nb=128, sched={"panel_cus": panel_cus})
    assert info == 4096 - len(zc)
    assert rep["fallbacks"] >= len(zc)
    check_parity(a, lu, ipiv, jpiv, info, marg)
