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

from helpers import golden, oracle, reconstruction_error, sha256, workloads, delayed_set


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


def test_planted_mistake_is_caught():
    """A mutant that breaks ties toward the highest row must fail the long-double check."""
    a = _tricky(12, 1, ties=True)
    u, ipu, jpu, infou = oracle.unblocked_ld(a)
    # mutant: same algorithm, last-index argmax for the first column
    col = np.abs(a[:, 0])
    mutant_first = int(np.flatnonzero(col == col.max())[-1])
    assert mutant_first != ipu[0]
