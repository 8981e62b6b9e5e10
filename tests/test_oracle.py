"""CPU checks that pin the oracle (oracle/dlu_oracle.c) before it is trusted as the checker.

Pins used (each chosen so that a plausible mistake fails one of them):
  * known answers from the reference's own tests (example.cu, example.cpp) with pivots derived
    by hand -- a wrong tie rule, a missed materialization or a wrong restart fails these;
  * LAPACK dgetrf: with no zero columns the algorithm is textbook partial pivoting, so ipiv must
    equal LAPACK's (0-based) and jpiv must be the identity;
  * an independent unblocked restatement in long double (no BLAS) on small inputs with zero
    columns and exact ties: the blocked restatement must choose the same pivots;
  * invariants: backward error, |L| <= 1, the zero-column delayed set at tau = 0;
  * golden vectors recorded from the oracle (reproducibility across hosts/BLAS builds).
"""
import numpy as np
import pytest

from helpers import delayed_set, dlu, golden, oracle, reconstruction_error, sha256, workloads


def _case_matrix(c):
    a = np.zeros((c["m"], c["n"]), order="F")
    if "dense_rows" in c:
        a[:, :] = np.array(c["dense_rows"])
    for i, j, v in c.get("entries", []):
        a[i, j] = v
    return a


@pytest.mark.parametrize("idx", range(4))
def test_known_answers_small(idx):
    c = golden("known_answers.json")["cases"][idx]
    a = _case_matrix(c)
    lu, ipiv, jpiv, info, _ = oracle.delayed_getrf(a)
    assert info == c["info"]
    assert ipiv.tolist() == c["ipiv"]
    assert jpiv.tolist() == c["jpiv"]
    tol = 1e-11 * max(1.0, np.abs(a).max()) * max(a.shape)  # delayed_lu_cuda/example.cu:101-106
    assert reconstruction_error(a, lu, ipiv, jpiv, info) <= tol


@pytest.mark.slow
def test_known_answer_example_cpp():
    c = golden("known_answers.json")["cases"][4]
    a = _case_matrix(c)
    lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a)
    assert info == 800
    assert ipiv.tolist() == c["ipiv"] and jpiv.tolist() == c["jpiv"]
    # delayed_lu/example.cpp:222-223: max error <= 100 eps * 800
    err = oracle.backward_error(a, lu, ipiv, jpiv, info)
    assert err >= 0 and err < 10


@pytest.mark.parametrize("m,n,seed", [(64, 64, 1), (200, 200, 2), (512, 512, 3), (300, 150, 4),
                                      (150, 300, 5), (1000, 1000, 6)])
def test_matches_lapack_without_zero_columns(m, n, seed):
    a = workloads.normal(m, n, seed)
    lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a, margins=True)
    l2, ipl, infol = oracle.lapack_getrf(a)
    assert info == min(m, n) and infol == min(m, n)
    assert ex["fallbacks"] == 0
    assert np.array_equal(jpiv, np.arange(min(m, n)))
    if not np.array_equal(ipiv, ipl):  # only a near-tie may separate them
        k = int(np.argmax(ipiv != ipl))
        assert ex["margins"][k] < 1e-10, f"step {k} differs with margin {ex['margins'][k]}"
    assert np.abs(lu - l2).max() <= 1e-12 * np.abs(l2).max()
    assert oracle.backward_error(a, lu, ipiv, jpiv, info) < 10


def _tricky(n, seed, zero_cols=(), ties=False, int_entries=True):
    rng = np.random.default_rng(seed)
    a = rng.integers(-4, 5, size=(n, n)).astype(np.float64) if int_entries else \
        rng.standard_normal((n, n))
    for j in zero_cols:
        a[:, j] = 0.0
    if ties:
        a[3, 0] = -a[:, 0].__abs__().max()  # equal magnitudes in one column
        a[5, 0] = a[:, 0].__abs__().max()
    return np.asfortranarray(a)


@pytest.mark.parametrize("seed", range(12))
def test_blocked_matches_unblocked_long_double(seed):
    n = 40
    rng = np.random.default_rng(100 + seed)
    zc = rng.choice(n, size=int(rng.integers(0, 8)), replace=False)
    a = _tricky(n, seed, zero_cols=zc, ties=seed % 2 == 0, int_entries=seed % 3 != 0)
    for nb in (1, 3, 8, 64):
        lu, ipiv, jpiv, info, _ = oracle.delayed_getrf(a, nb=nb)
        u, ipu, jpu, infou = oracle.unblocked_ld(a)
        assert info == infou
        assert ipiv.tolist() == ipu.tolist(), f"nb={nb}"
        assert jpiv.tolist() == jpu.tolist(), f"nb={nb}"
        assert np.abs(lu - u).max() <= 1e-12 * max(1.0, np.abs(u).max())


@pytest.mark.parametrize("n,seed", [(512, 3), (1024, 5)])
def test_zero_column_invariant(n, seed):
    a, zc, tc, off = workloads.delay_stress(n, seed)
    lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a)
    assert info == n - len(zc)
    assert np.array_equal(delayed_set(jpiv, info, n), zc)
    assert np.all(lu[info:, info:] == 0.0)  # the Schur complement of zero columns stays exactly 0
    assert ex["fallbacks"] >= 1
    assert oracle.backward_error(a, lu, ipiv, jpiv, info) < 10
    assert np.abs(np.tril(lu[:, :info], -1)).max() <= 1.0


def test_golden_vectors():
    d = golden("oracle_vectors.json")
    for c in d["cases"]:
        if c["kind"] == "normal":
            a = workloads.normal(c["m"], c["n"], c["seed"])
        else:
            a = workloads.delay_stress(c["n"], c["seed"])[0]
        assert sha256(a) == c["sha256"], c["name"]
        lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a)
        assert info == c["info"], c["name"]
        assert ipiv.tolist() == c["ipiv"], c["name"]
        assert jpiv.tolist() == c["jpiv"], c["name"]
        assert ex["fallbacks"] == c["fallbacks"], c["name"]


def test_argument_rules():
    a = np.ones((3, 3), order="F")
    # delayed_lu/delayed_dgetrf.cpp:132-142: info = -1 on invalid arguments
    import ctypes
    L = oracle.lib()
    info = ctypes.c_int(5)
    ip = np.zeros(3, dtype=np.int32)
    L.dlu_oracle_delayed_getrf(3, 3, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 2,
                               ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               ctypes.byref(info), 64, None, None)
    assert info.value == -1
    lu, ipiv, jpiv, info, _ = oracle.delayed_getrf(np.zeros((0, 5)))
    assert info == 0
    lu, ipiv, jpiv, info, _ = oracle.delayed_getrf(np.zeros((4, 4)))
    assert info == 0 and ipiv.tolist() == [0, 1, 2, 3]


def test_planted_mistake_is_caught(tmp_path):
    """A mutant oracle, built from the same source with the column-pivot tie rule flipped to the
    highest row (DLU_ORACLE_MUTANT_TIES), must disagree with the long-double pin on an input with
    exact ties, while the real build agrees: the pin can see a wrong tie rule."""
    path = oracle.build(force=True, out=str(tmp_path / "libdlu_oracle_mutant.so"),
                        defines=["DLU_ORACLE_MUTANT_TIES"])
    mutant = oracle.load_build(path)
    a = _tricky(12, 1, ties=True)
    _, ipu, jpu, infou = oracle.unblocked_ld(a)
    _, ipm, jpm, infom, _ = oracle.delayed_getrf(a, library=mutant)
    _, ipr, jpr, infor, _ = oracle.delayed_getrf(a)
    assert ipr.tolist() == ipu.tolist() and jpr.tolist() == jpu.tolist()
    assert ipm.tolist() != ipu.tolist()
    ties = np.flatnonzero(np.abs(a[:, 0]) == np.abs(a[:, 0]).max())
    assert len(ties) >= 2 and ipu[0] == ties[0] and ipm[0] == ties[-1]


def _frobenius_backward_error(a, lu, ipiv, jpiv, info):
    """||P A Q - Lbar Ubar||_F / (max(m,n) eps ||A||_F), written independently in numpy."""
    m, n = a.shape
    lo = np.eye(m)
    up = np.zeros((m, n))
    for k in range(info):
        lo[k + 1:, k] = lu[k + 1:, k]
        up[k, k:] = lu[k, k:]
    up[info:, info:] = lu[info:, info:]
    rows = oracle.replay(ipiv, info, m)
    cols = oracle.replay(jpiv, info, n)
    d = lo @ up - a[np.ix_(rows, cols)]
    return np.linalg.norm(d) / (max(m, n) * np.finfo(float).eps * np.linalg.norm(a))


@pytest.mark.parametrize("case", ["normal", "rect", "stress"])
def test_backward_error_two_sided(case):
    """dlu_oracle_backward_error agrees with an independent numpy reconstruction, and reports a
    large value for a corrupted factor and for a wrong pivot history."""
    if case == "normal":
        a = workloads.normal(300, 300, 41)
    elif case == "rect":
        a = workloads.normal(260, 190, 42)
    else:
        a = workloads.delay_stress(256, 43)[0]
    lu, ipiv, jpiv, info, _ = oracle.delayed_getrf(a)
    got = oracle.backward_error(a, lu, ipiv, jpiv, info)
    ref = _frobenius_backward_error(a, lu, ipiv, jpiv, info)
    assert 0 <= got < 10
    # correct factors leave a residual at rounding level, which itself depends on the order in
    # which L Ubar is formed: the two agree to a fraction of one unit (n eps ||A||_F)
    assert abs(got - ref) <= 0.1 * max(got, ref) + 0.05
    bad = lu.copy()
    bad[info // 2 + 3, info // 3] += 1e-8 * np.abs(lu).max()  # one multiplier off by 1e-8
    got_bad = oracle.backward_error(a, bad, ipiv, jpiv, info)
    assert got_bad > 1e3
    assert abs(got_bad - _frobenius_backward_error(a, bad, ipiv, jpiv, info)) <= 1e-6 * got_bad
    wrong = ipiv.copy()
    k = int(np.argmax(wrong != np.arange(len(wrong))))
    wrong[k] = k  # drop one row interchange
    assert oracle.backward_error(a, lu, wrong, jpiv, info) > 1e6


# ---- the column-delay variant (delayed_lu/delayed_dgetrf_column.cpp) -----------------------

def test_column_variant_known_answer_example_cpp():
    """example.cpp's 4096 x 4096 rank-800 matrix (diag(2..801) = 800..1): steps 0 and 1 find
    their columns zero and take the rightmost nonzero columns, 801 then 800; the rest are
    column pivots; step 800 finds every remaining column zero (info = 800)."""
    a = workloads.diag_rank(4096, 800)
    lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a, column=True)
    want = [801, 800] + list(range(2, 800)) + list(range(800, 4096))
    assert info == 800
    assert ipiv.tolist() == want and jpiv.tolist() == want
    assert ex["fallbacks"] == 3
    assert oracle.backward_error(a, lu, ipiv, jpiv, info) < 10


@pytest.mark.parametrize("seed", range(12))
def test_column_blocked_matches_unblocked_long_double(seed):
    n = 40
    rng = np.random.default_rng(300 + seed)
    zc = rng.choice(n, size=int(rng.integers(1, 9)), replace=False)
    a = _tricky(n, seed + 50, zero_cols=zc, ties=seed % 2 == 0, int_entries=seed % 3 != 0)
    u, ipu, jpu, infou = oracle.unblocked_ld(a, column=True)
    for nb in (1, 3, 8, 64):
        lu, ipiv, jpiv, info, _ = oracle.delayed_getrf(a, nb=nb, column=True)
        assert info == infou
        assert ipiv.tolist() == ipu.tolist(), f"nb={nb}"
        assert jpiv.tolist() == jpu.tolist(), f"nb={nb}"
        assert np.abs(lu - u).max() <= 1e-12 * max(1.0, np.abs(u).max())


def test_column_variant_takes_rightmost_column():
    """Column 0 zero: the full-pivot variant takes the global maximum's column, the column
    variant the last column (its largest |a|, lowest row on ties)."""
    n = 30
    a = _tricky(n, 7, int_entries=True)
    a[:, 0] = 0.0
    a[:, n - 1] = 0.0
    a[4, n - 1] = -2.0
    a[9, n - 1] = 2.0
    a[17, 11] = 50.0
    _, ipf, jpf, _, _ = oracle.delayed_getrf(a)
    _, ipc, jpc, _, _ = oracle.delayed_getrf(a, column=True)
    assert (ipf[0], jpf[0]) == (17, 11)
    assert (ipc[0], jpc[0]) == (4, n - 1)


@pytest.mark.parametrize("n,seed", [(512, 3), (1024, 5)])
def test_column_zero_column_invariant(n, seed):
    """At tau = 0 the delayed set is the zero-column set for this variant too, and it is reached
    with one fallback per zero column met (each zero column goes to the right end)."""
    a, zc, tc, off = workloads.delay_stress(n, seed)
    lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a, column=True)
    assert info == n - len(zc)
    assert np.array_equal(delayed_set(jpiv, info, n), zc)
    assert np.all(lu[info:, info:] == 0.0)
    assert oracle.backward_error(a, lu, ipiv, jpiv, info) < 10
    _, _, _, _, exf = oracle.delayed_getrf(a)
    assert ex["fallbacks"] <= len(zc) + 1 < exf["fallbacks"]


# ---- frontal-matrix Step 2 and the extend-add index logic -----------------------------------

def _front(nf, ns, seed, zero_fs_cols=()):
    a = workloads.normal(nf, nf, seed)
    a[:ns, list(zero_fs_cols)] = 0.0  # zero in the fully summed rows: delayed at tau = 0
    a[ns:, list(zero_fs_cols)] = 0.0
    return a


@pytest.mark.parametrize("nf,ns,zeros", [(60, 40, ()), (60, 40, (3, 17)), (90, 90, (5,)),
                                         (50, 20, (0, 19))])
def test_front_oracle_reconstructs_and_gives_the_schur_complement(nf, ns, zeros):
    """The front oracle's output is a block LU of the front with the fully summed interchanges
    (P F Q = Lbar Ubar, backward error <= 10), and its contribution block equals the Schur
    complement S = F[R, C] - F[R, Ec] F[Er, Ec]^-1 F[Er, C] computed directly (Er / Ec: the
    eliminated rows / columns, R / C the rest)."""
    a = _front(nf, ns, nf * 7 + ns, zeros)
    f, ipiv, jpiv, e, = oracle.front_factor(a, ns)
    assert e == ns - len(zeros)
    rows = np.arange(nf)
    cols = np.arange(nf)
    rows[:ns] = oracle.replay(ipiv, e, ns)
    cols[:ns] = oracle.replay(jpiv, e, ns)
    er, ec = rows[:e], cols[:e]
    rr, cc = rows[e:], cols[e:]
    s = a[np.ix_(rr, cc)] - a[np.ix_(rr, ec)] @ np.linalg.solve(a[np.ix_(er, ec)], a[np.ix_(er, cc)])
    assert np.abs(f[e:, e:] - s).max() <= 1e-10 * max(1.0, np.abs(s).max())
    # full reconstruction: the front as a compact factorization with e eliminated steps whose
    # interchanges are those of the fully summed block
    ip = np.concatenate([ipiv, np.arange(ns, nf)]).astype(np.int32)
    jp = np.concatenate([jpiv, np.arange(ns, nf)]).astype(np.int32)
    assert oracle.backward_error(a, f, ip, jp, e) <= 10.0


def test_front_oracle_without_update_part_is_delayed_getrf():
    a = _front(70, 70, 3, (4,))
    f, ipiv, jpiv, e = oracle.front_factor(a, 70)
    lu, ip2, jp2, e2, _ = oracle.delayed_getrf(a)
    assert e == e2 and np.array_equal(ipiv, ip2) and np.array_equal(jpiv, jp2)
    assert np.array_equal(f, lu)


def test_contribution_indices_and_parent_positions():
    # a front over global variables [10, 11, 12, 13 | 20, 21] (4 fully summed, 2 update)
    row_ids, col_ids = [10, 11, 12, 13, 20, 21], [10, 11, 12, 13, 20, 21]
    ipiv = np.array([2, 1, 2, 3], dtype=np.int32)  # step 0 swaps rows 0 and 2
    jpiv = np.array([3, 1, 2, 3], dtype=np.int32)  # step 0 swaps cols 0 and 3
    r, c = dlu.contribution_indices(row_ids, col_ids, 4, ipiv, jpiv, 2)
    assert r.tolist() == [10, 13, 20, 21]  # rows after the swap: 12, 11, 10, 13
    assert c.tolist() == [12, 10, 20, 21]  # cols after the swap: 13, 11, 12, 10
    ids, pos = dlu.parent_positions([20, 21, 30], [10, 13, 20, 21])
    assert ids.tolist() == [20, 21, 30, 10, 13] and pos.tolist() == [3, 4, 0, 1]


# ---- delayed Bunch-Kaufman LDL^T (delayed-bunch-kaufman/) ------------------------------------

def _sym(n, seed, zero_nodes=()):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n))
    s = np.asfortranarray(x + x.T)
    z = list(zero_nodes)
    s[z, :] = 0.0
    s[:, z] = 0.0
    return s


def _lapack_dsytrf_pivots(s):
    """LAPACK DSYTRF (lower)'s interchanges in the delayed format: perm by applying each step's
    symmetric interchange in order, piv_size 1 / 2 0 (delayed-bunch-kaufman/lapack_dsytrf.cpp
    converts the same way through DSYCONV)."""
    import scipy.linalg.lapack as la
    n = s.shape[0]
    _, ipiv, info = la.dsytrf(s, lower=1)
    perm, piv, k = np.arange(n), np.zeros(n, dtype=int), 0
    while k < n:
        if ipiv[k] > 0:
            p = ipiv[k] - 1
            perm[[k, p]] = perm[[p, k]]
            piv[k] = 1
            k += 1
        else:
            p = -ipiv[k] - 1
            perm[[k + 1, p]] = perm[[p, k + 1]]
            piv[k], piv[k + 1] = 2, 0
            k += 2
    return perm, piv, info


def _example_cpp_sytrf():
    """delayed-bunch-kaufman/example.cpp:248-266: n = 4096, A(3,2) = A(2,3) = 2 (a 2x2 indefinite
    pivot), A(4,4) .. A(9,9) = 6 .. 1 (six 1x1 pivots), every other entry zero."""
    n = 4096
    a = np.zeros((n, n), order="F")
    a[3, 2] = a[2, 3] = 2.0
    for k in range(6):
        a[k + 4, k + 4] = 6.0 - k
    return a


@pytest.mark.parametrize("mode", ["blocked", "unblocked", "ld"])
def test_sytrf_known_answer_example_cpp(mode):
    """Derived by hand: node 0 is zero and is deferred to position 4095; each node then brought
    to position 0 (4095, 4094, ..., 10) is zero and deferred to the front of the suffix, until
    node 9 arrives (1x1, d = 1).  Node 1 (zero) is deferred to position 8 and node 8 takes its
    place (1x1, d = 2); nodes 2, 3 have zero diagonals and A(3,2) = 2 (colmax = rowmax = 2, the
    2x2 block [[0, 2], [2, 0]] is stable); nodes 4..7 are 1x1 (6, 5, 4, 3).  info = 8, as
    example.cpp's check expects (rank 8)."""
    a = _example_cpp_sytrf()
    f, perm, piv, info = oracle.delayed_sytrf(a, blocked=mode == "blocked", ld=mode == "ld")
    n = a.shape[0]
    assert info == 8
    assert perm.tolist() == [9, 8, 2, 3, 4, 5, 6, 7, 1] + list(range(10, n)) + [0]
    assert piv.tolist() == [1, 1, 2, 0, 1, 1, 1, 1] + [0] * (n - 8)
    assert [f[0, 0], f[1, 1], f[2, 2], f[3, 2], f[3, 3]] == [1.0, 2.0, 0.0, 2.0, 0.0]
    assert [f[k, k] for k in range(4, 8)] == [6.0, 5.0, 4.0, 3.0]
    assert np.count_nonzero(np.tril(f, -1)[:, :8]) == 1  # only the 2x2 block's A(3,2): L = 0
    assert not f[8:, 8:].any()
    assert oracle.sytrf_reconstruction_error(a, f, perm, piv, info) == 0.0


@pytest.mark.parametrize("n,seed", [(100, 1), (300, 2), (700, 3), (1500, 4)])
def test_sytrf_matches_lapack_dsytrf(n, seed):
    """Without delays the rules are LAPACK's Bunch-Kaufman (alpha = (1 + sqrt 17) / 8, the same
    three tests), so the interchanges and block sizes must equal DSYTRF's."""
    s = _sym(n, seed)
    perm_l, piv_l, info_l = _lapack_dsytrf_pivots(s)
    f, perm, piv, info = oracle.delayed_sytrf(s)
    assert info_l == 0 and info == n
    assert np.array_equal(perm, perm_l) and np.array_equal(piv, piv_l)
    assert (piv == 2).sum() > 0  # the case exercises 2x2 pivots
    assert oracle.sytrf_reconstruction_error(s, f, perm, piv, info) < 1e-13


@pytest.mark.parametrize("n,seed,nz", [(60, 5, 6), (200, 6, 20), (333, 7, 40), (150, 8, 0)])
def test_sytrf_restatements_agree(n, seed, nz):
    """The blocked restatement at panel widths 64, 16 and 2, the unblocked one, and the
    independent long-double one with no BLAS choose the same pivots and the same delayed set;
    every factorization reconstructs P^T A P."""
    z = np.random.default_rng(seed).choice(n, size=nz, replace=False)
    s = _sym(n, seed, z)
    ref = oracle.delayed_sytrf(s, ld=True)
    for kw in ({"nb": 64}, {"nb": 16}, {"nb": 2}, {"blocked": False}):
        f, perm, piv, info = oracle.delayed_sytrf(s, **kw)
        assert info == ref[3] == n - nz, kw
        assert np.array_equal(perm, ref[1]) and np.array_equal(piv, ref[2]), kw
        assert oracle.sytrf_reconstruction_error(s, f, perm, piv, info) < 1e-13, kw
    assert sorted(ref[1][ref[3]:].tolist()) == sorted(z.tolist())  # the delayed set


def test_sytrf_leaves_the_upper_triangle_alone():
    s = _sym(90, 9)
    s[np.triu_indices(90, 1)] = 7.25  # not the symmetric values: never read, never written
    f, perm, piv, info = oracle.delayed_sytrf(s)
    assert np.all(f[np.triu_indices(90, 2)] == 7.25)
    g = oracle.delayed_sytrf(np.tril(s) + np.tril(s, -1).T)
    assert np.array_equal(perm, g[1]) and np.array_equal(piv, g[2])


def test_sytrf_unstable_2x2_is_deferred():
    """A candidate pair whose 2x2 block is singular to working precision ([[1, 2], [2, 4]],
    with no diagonal large enough for a 1x1 pivot) is not accepted: node 0 goes to the delayed
    suffix."""
    a = np.zeros((3, 3), order="F")
    a[0, 0], a[1, 0], a[1, 1] = 1.0, 2.0, 4.0
    a[2, 2] = 5.0
    for kw in ({}, {"blocked": False}, {"ld": True}):
        f, perm, piv, info = oracle.delayed_sytrf(a, **kw)
        assert info == 2 and perm[2] == 0, (kw, perm, info)


def test_sytrf_planted_mistake_is_caught(tmp_path):
    """A mutant built with DLU_ORACLE_MUTANT_BK_NOSWAP takes the candidate's diagonal as a 1x1
    pivot without bringing the candidate to k: the long-double pin must disagree with it on an
    input where that rule fires, and agree with the real build."""
    path = oracle.build(force=True, out=str(tmp_path / "libdlu_oracle_bk_mutant.so"),
                        defines=["DLU_ORACLE_MUTANT_BK_NOSWAP"])
    mutant = oracle.load_build(path)
    a = np.zeros((3, 3), order="F")
    # |a00| < alpha |a10|, |a11| >= alpha * rowmax; node 0's Schur complement after node 1,
    # 0.2 - 1 / 10, does not cancel
    a[0, 0], a[1, 0], a[1, 1], a[2, 2] = 0.2, 1.0, 10.0, 3.0
    ref = oracle.delayed_sytrf(a, ld=True)
    real = oracle.delayed_sytrf(a)
    bad = oracle.delayed_sytrf(a, library=mutant)
    assert ref[1].tolist() == real[1].tolist() and ref[1][0] == 1
    assert bad[1].tolist() != ref[1].tolist()
