"""GPU checks of the frontal-matrix steps (the design note's multifrontal use of the delayed
suffix, doc/（非）对称稠密矩阵的延迟主元LU方法.md:228-440), through the C ABI:

  * dlu_front_factor_device against the front oracle (oracle.front_factor): the fully summed
    block's ipiv / jpiv / info bit-exact, the whole front within 1e-12 relative (Frobenius);
  * dlu_extend_add_device against numpy's scatter-add, several children into one parent;
  * a two-level assembly: a child front with delayed variables, its contribution block
    extend-added into the parent (the delayed variables join the parent's fully summed set), the
    assembled parent equal to the global Schur complement computed directly, and the parent's
    factorization eliminating the child's delayed variables.
"""
import numpy as np
import pytest

from helpers import dlu, oracle, workloads

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dlu.lib()
    return torch


def _front(nf, ns, seed, zeros=()):
    a = workloads.normal(nf, nf, seed)
    a[:, list(zeros)] = 0.0  # zero fully summed columns: delayed at tau = 0
    return a


def gpu_front(a, ns):
    torch = _torch()
    nf = a.shape[0]
    t, ldf = dlu.to_device(a)
    st, ipiv, jpiv, info = dlu.front_factor_device(nf, ns, t.data_ptr(), ldf)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS, dlu.status_string(st)
    return dlu.from_device(t, nf), ipiv, jpiv, info


@pytest.mark.parametrize("nf,ns,zeros", [(300, 200, ()), (300, 200, (7, 150)), (700, 500, (0,)),
                                         (400, 400, (10,)), (257, 1, ()), (1200, 900, (450, 451)),
                                         (520, 130, (129,))])
def test_front_factor_matches_oracle(nf, ns, zeros):
    a = _front(nf, ns, nf * 3 + ns, zeros)
    f, ipiv, jpiv, info = gpu_front(a, ns)
    fc, ip_c, jp_c, info_c = oracle.front_factor(a, ns)
    assert info == info_c == ns - len(zeros)
    assert np.array_equal(ipiv, ip_c) and np.array_equal(jpiv, jp_c)
    assert np.linalg.norm(f - fc) <= 1e-12 * np.linalg.norm(fc)


def test_front_factor_empty_fully_summed_block():
    a = _front(50, 0, 5)
    f, ipiv, jpiv, info = gpu_front(a, 0)
    assert info == 0 and np.array_equal(f, a)


def test_extend_add_matches_numpy():
    torch = _torch()
    rng = np.random.default_rng(8)
    n_parent = 500
    parent = rng.standard_normal((n_parent, n_parent))
    ref = parent.copy()
    tp, ldp = dlu.to_device(parent)
    keep = []
    for rows, cols in ((120, 90), (300, 260), (1, 7)):
        child = rng.standard_normal((rows, cols))
        mr = rng.choice(n_parent, size=rows, replace=False).astype(np.int32)
        mc = rng.choice(n_parent, size=cols, replace=False).astype(np.int32)
        np.add.at(ref, (mr[:, None], mc[None, :]), child)
        tc, ldc = dlu.to_device(child)
        dmr = torch.from_numpy(mr).to("cuda")
        dmc = torch.from_numpy(mc).to("cuda")
        keep += [tc, dmr, dmc]
        torch.cuda.synchronize()
        assert dlu.extend_add_device(rows, cols, tc.data_ptr(), ldc, dmr.data_ptr(),
                                     dmc.data_ptr(), tp.data_ptr(), ldp) == dlu.Status.SUCCESS
    torch.cuda.synchronize()
    assert np.abs(dlu.from_device(tp, n_parent) - ref).max() <= 1e-13 * np.abs(ref).max()


def test_two_level_assembly_eliminates_delayed_variables():
    """Global variables 0..N-1.  The child front holds C = [0, c) fully summed plus an update set
    Uc shared with the parent (the root), which holds P = [c, N).  Within the child's rows,
    column 5 is exactly twice column 0 (powers of two, so the elimination of column 0 leaves it
    exactly zero there): the child delays it (tau = 0) although its Uc rows are not zero.  Its
    contribution block goes to the parent by extend-add, the delayed row and column join the
    parent's fully summed set, the assembled parent equals the global Schur complement computed
    directly, and the parent's factorization eliminates everything."""
    torch = _torch()
    rng = np.random.default_rng(21)
    N, c = 600, 200
    P = np.arange(c, N)
    uc = np.sort(rng.choice(P, size=120, replace=False))
    C = np.arange(c)
    a = np.zeros((N, N))
    a[c:, c:] = rng.standard_normal((N - c, N - c))      # the parent's own entries
    a[:c, :c] = rng.standard_normal((c, c))              # the child's fully summed block
    a[np.ix_(C, uc)] = rng.standard_normal((c, len(uc)))
    a[np.ix_(uc, C)] = rng.standard_normal((len(uc), c))
    a[:c, 0] = rng.choice([-1.0, 1.0], size=c) * 2.0 ** rng.integers(-3, 4, size=c)
    a[17, 0] = 16.0                                       # the unique largest: step 0's pivot
    a[:c, 5] = 2.0 * a[:c, 0]
    a = np.asfortranarray(a)
    ids = np.concatenate([C, uc])
    front = a[np.ix_(ids, ids)].copy()
    front[c:, c:] = 0.0  # the Uc x Uc entries belong to the parent's assembly
    f, ipiv, jpiv, e = gpu_front(np.asfortranarray(front), c)
    assert e == c - 1  # one delayed variable
    cb_rows, cb_cols = dlu.contribution_indices(ids, ids, c, ipiv, jpiv, e)
    assert cb_cols[0] == 5 and len(cb_rows) == len(ids) - e
    prow_ids, map_r = dlu.parent_positions(P, cb_rows)
    pcol_ids, map_c = dlu.parent_positions(P, cb_cols)
    npar = len(prow_ids)
    assert npar == len(pcol_ids) == N - c + 1 and pcol_ids[-1] == 5
    parent = np.zeros((npar, npar))
    parent[:N - c, :N - c] = a[c:, c:]  # the delayed row and column have no entries of their own
    tpar, ldp = dlu.to_device(np.asfortranarray(parent))
    cb = np.asfortranarray(f[e:, e:])
    tcb, ldc = dlu.to_device(cb)
    dmr = torch.from_numpy(map_r).to("cuda")
    dmc = torch.from_numpy(map_c).to("cuda")
    torch.cuda.synchronize()
    assert dlu.extend_add_device(cb.shape[0], cb.shape[1], tcb.data_ptr(), ldc, dmr.data_ptr(),
                                 dmc.data_ptr(), tpar.data_ptr(), ldp) == dlu.Status.SUCCESS
    torch.cuda.synchronize()
    assembled = dlu.from_device(tpar, npar)
    er = ids[oracle.replay(ipiv, e, c)[:e]]
    ec = ids[oracle.replay(jpiv, e, c)[:e]]
    s = a[np.ix_(prow_ids, pcol_ids)] - a[np.ix_(prow_ids, ec)] @ np.linalg.solve(
        a[np.ix_(er, ec)], a[np.ix_(er, pcol_ids)])
    assert np.abs(assembled - s).max() <= 1e-9 * np.abs(s).max()
    assert np.abs(assembled[:, -1]).max() > 0.0  # the delayed column is no longer zero
    st, ipp, jpp, infop = dlu.getrf_device(npar, npar, tpar.data_ptr(), ldp)[:4]
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS and infop == npar
    lu = dlu.from_device(tpar, npar)
    assert oracle.backward_error(assembled, lu, ipp, jpp, infop) <= 10.0
