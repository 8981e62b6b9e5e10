"""GPU checks of the batched small-matrix kernel and the solve service (C ABI).

  * batched factor + solve: info / ipiv / jpiv / delayed set equal to the oracle's (the reference
    algorithm on the CPU) item by item -- N(0,1), zero columns (full-pivot fallbacks), exact ties,
    an all-zero matrix; factors within 1e-12 normwise; solutions by the numpy backward error;
  * a result does not depend on the batch it ran in (bitwise), and solve-only items reuse the
    factors exactly;
  * the service: each request's solution is bitwise the same submitted alone and in a mix of
    small and large, fresh and reused requests; evicted factors are rebuilt; the mixed mode
    reaches the FP64 backward error; refinement sweeps replayed from HIP graphs give the same
    bits as sweeps launched kernel by kernel.
"""
import numpy as np
import pytest

from helpers import delayed_set, dlu, oracle, workloads

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    dlu.lib()
    return torch


class Dev:
    """A column-major FP64 device matrix (storage tensor rows are columns)."""

    def __init__(self, a):
        torch = _torch()
        self.m, self.n = a.shape
        self.t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).to("cuda")

    @property
    def ptr(self):
        return self.t.data_ptr()

    def host(self):
        return np.asfortranarray(self.t.cpu().numpy().T)


def _ints(k):
    torch = _torch()
    return torch.zeros(max(k, 1), dtype=torch.int32, device="cuda")


def _matrices():
    rng = np.random.default_rng(5)
    mats = []
    for n in (1, 2, 17, 64, 65, 100, 128):
        mats.append(workloads.normal(n, n, 300 + n))
    a = workloads.normal(96, 96, 7)
    a[:, [3, 40, 41]] = 0.0  # zero columns: fallbacks, then a delayed suffix
    mats.append(a)
    a = workloads.normal(80, 80, 8)
    a[:, 10] = 0.0
    a[5, 10] = 3.0
    a[70, 10] = -3.0  # exact tie inside a column: the lower row wins
    mats.append(a)
    mats.append(np.zeros((33, 33), order="F"))
    b = np.outer(2.0 ** rng.integers(-3, 4, 50), 2.0 ** rng.integers(-3, 4, 50))
    mats.append(np.asfortranarray(b))  # rank one with exact ties everywhere
    return mats


def _run_batch(mats, rhs, order=None):
    torch = _torch()
    order = list(range(len(mats))) if order is None else order
    bufs, items = [], []
    for i in order:
        a = mats[i]
        n = a.shape[0]
        da, dlu_ = Dev(a), Dev(np.zeros_like(a))
        db, dx = Dev(rhs[i]), Dev(np.zeros_like(rhs[i]))
        piv, info = _ints(2 * n), _ints(1)
        it = dlu.BatchItem(a=da.ptr, lda=n, lu=dlu_.ptr, ldlu=n, ipiv=piv.data_ptr(),
                           jpiv=piv.data_ptr() + 4 * n, info=info.data_ptr(), info_in=0, n=n,
                           nrhs=rhs[i].shape[1], b=db.ptr, ldb=n, x=dx.ptr, ldx=n, solve_only=0)
        items.append(it)
        bufs.append((i, da, dlu_, db, dx, piv, info))
    st = dlu.batch_gesv_device(items)
    torch.cuda.synchronize()
    assert st == dlu.Status.SUCCESS
    out = {}
    for i, da, dl, db, dx, piv, info in bufs:
        n = mats[i].shape[0]
        p = piv.cpu().numpy()
        out[i] = dict(lu=dl.host(), ipiv=p[:n].copy(), jpiv=p[n:2 * n].copy(),
                      info=int(info.cpu().numpy()[0]), x=dx.host(), bufs=(dl, piv, info))
    return out


def test_batch_matches_oracle_item_by_item():
    mats = _matrices()
    rhs = [np.asfortranarray(np.random.default_rng(i).standard_normal((a.shape[0], 1 + i % 8)))
           for i, a in enumerate(mats)]
    res = _run_batch(mats, rhs)
    for i, a in enumerate(mats):
        r = res[i]
        n = a.shape[0]
        lu_c, ip_c, jp_c, info_c, _ = oracle.delayed_getrf(a, nb=64)
        assert r["info"] == info_c, i
        assert np.array_equal(r["ipiv"], ip_c) and np.array_equal(r["jpiv"], jp_c), i
        assert np.array_equal(delayed_set(r["jpiv"], r["info"], n),
                              delayed_set(jp_c, info_c, n))
        assert np.linalg.norm(r["lu"] - lu_c) <= 1e-12 * max(np.linalg.norm(lu_c), 1e-300)
        if r["info"] == n:
            x, b = r["x"], rhs[i]
            res_ = np.abs(b - a @ x).max(axis=0)
            den = np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0)
            assert (res_ / den).max() <= 1e-14, i


def test_batch_result_does_not_depend_on_the_batch():
    mats = _matrices()
    rhs = [np.asfortranarray(np.random.default_rng(50 + i).standard_normal((a.shape[0], 2)))
           for i, a in enumerate(mats)]
    together = _run_batch(mats, rhs, order=list(reversed(range(len(mats)))))
    for i in (0, 4, 6, 7):
        alone = _run_batch(mats, rhs, order=[i])[i]
        assert np.array_equal(alone["lu"], together[i]["lu"])
        assert np.array_equal(alone["x"], together[i]["x"])
        assert np.array_equal(alone["ipiv"], together[i]["ipiv"])


def test_batch_solve_only_reuses_factors():
    torch = _torch()
    a = workloads.normal(90, 90, 4)
    rhs1 = np.asfortranarray(np.random.default_rng(1).standard_normal((90, 3)))
    r = _run_batch([a], [rhs1])[0]
    dl, piv, info = r["bufs"]
    rhs2 = np.asfortranarray(np.random.default_rng(2).standard_normal((90, 5)))
    db, dx = Dev(rhs2), Dev(np.zeros_like(rhs2))
    it = dlu.BatchItem(a=dl.ptr, lda=90, lu=0, ldlu=90, ipiv=piv.data_ptr(),
                       jpiv=piv.data_ptr() + 4 * 90, info=0, info_in=r["info"], n=90, nrhs=5,
                       b=db.ptr, ldb=90, x=dx.ptr, ldx=90, solve_only=1)
    assert dlu.batch_gesv_device([it]) == dlu.Status.SUCCESS
    torch.cuda.synchronize()
    x = dx.host()
    assert np.linalg.norm(x - np.linalg.solve(a, rhs2)) <= 1e-12 * np.linalg.norm(x)


def test_batch_rejects_bad_items():
    it = dlu.BatchItem(a=0, lda=10, n=10, nrhs=1)
    assert dlu.batch_gesv_device([it]) == dlu.Status.INVALID_ARGUMENT
    it = dlu.BatchItem(a=1, lda=200, n=200, nrhs=1)  # above the batch size limit
    assert dlu.batch_gesv_device([it]) == dlu.Status.INVALID_ARGUMENT


def _requests(seed, count, sizes):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(sizes[i % len(sizes)])
        nrhs = int(rng.integers(1, 9))
        reuse = -1
        if i >= 4 and rng.random() < 0.3:
            j = int(rng.integers(0, i))
            if reqs[j]["keep"]:
                reuse = j
                n = reqs[j]["n"]
        a = reqs[reuse]["a"] if reuse >= 0 else workloads.normal(n, n, 1000 * seed + i)
        b = np.asfortranarray(rng.standard_normal((n, nrhs)))
        reqs.append(dict(id=i, n=n, nrhs=nrhs, a=a, b=b, reuse=reuse, keep=bool(rng.random() < 0.6)))
    return reqs


def _serve(reqs, one_at_a_time, **cfg):
    torch = _torch()
    srv = dlu.Server(**cfg)
    dev = {}
    for r in reqs:
        da = dev[r["reuse"]][0] if r["reuse"] >= 0 else Dev(r["a"])
        db, dx = Dev(r["b"]), Dev(np.zeros_like(r["b"]))
        dev[r["id"]] = (da, db, dx)
        st = srv.submit(r["id"], r["n"], r["nrhs"], da.ptr, r["n"], db.ptr, r["n"], dx.ptr,
                        r["n"], factor_of=r["reuse"], keep=r["keep"])
        assert st == dlu.Status.SUCCESS
        if one_at_a_time:
            srv.drain()
    srv.drain()
    comps = {c["id"]: c for c in srv.poll(100000)}
    cutoff = srv.cutoff
    srv.close()
    torch.cuda.synchronize()
    xs = {i: d[2].host() for i, d in dev.items()}
    return comps, xs, cutoff


def test_server_result_is_the_same_alone_and_in_the_mix():
    reqs = _requests(3, 40, [40, 128, 300, 96, 700, 64, 1000, 17])
    c1, x1, cut = _serve(reqs, True, batch_cutoff=-1)  # cutoff timed by the service
    c2, x2, _ = _serve(reqs, False, batch_cutoff=cut)
    assert cut >= 32
    assert len(c1) == len(c2) == len(reqs)
    for r in reqs:
        i = r["id"]
        assert c1[i]["status"] == 0 and c2[i]["status"] == 0, (i, c1[i], c2[i])
        assert np.array_equal(x1[i], x2[i]), i
        ref = np.linalg.solve(r["a"], r["b"])
        assert np.linalg.norm(x1[i] - ref) <= 1e-10 * np.linalg.norm(ref)
        assert c1[i]["path"] == (0 if r["n"] <= cut else 1)
    assert any(c["reused"] for c in c2.values())


def test_server_rebuilds_evicted_factors():
    reqs = _requests(4, 24, [120, 200, 64])
    comps, xs, _ = _serve(reqs, False, budget_gib=0.0006, batch_cutoff=128)  # ~600 KiB
    for r in reqs:
        assert comps[r["id"]]["status"] == 0
        ref = np.linalg.solve(r["a"], r["b"])
        assert np.linalg.norm(xs[r["id"]] - ref) <= 1e-10 * np.linalg.norm(ref)


def test_server_mixed_mode_reaches_fp64_backward_error():
    torch = _torch()
    n = 600
    a = workloads.conditioned(n, 31)
    b = np.asfortranarray(np.random.default_rng(32).standard_normal((n, 3)))
    srv = dlu.Server(mixed=True, batch_cutoff=64)
    da, db, dx = Dev(a), Dev(b), Dev(np.zeros_like(b))
    assert srv.submit(1, n, 3, da.ptr, n, db.ptr, n, dx.ptr, n) == dlu.Status.SUCCESS
    srv.drain()
    c = srv.poll()[0]
    srv.close()
    torch.cuda.synchronize()
    assert c["status"] == 0 and c["path"] == 2 and c["backward_error"] <= 1e-13
    x = dx.host()
    r = np.abs(b - a @ x).max(axis=0)
    den = np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0)
    assert (r / den).max() <= 1e-13


def test_server_small_max_batch_with_tuned_cutoff():
    """max_batch below the start-up cutoff timing's batch (128 matrices) with the cutoff timed at
    creation: the timing's item arrays must still fit, and the service must work."""
    reqs = _requests(6, 30, [40, 64, 100, 300])
    comps, xs, cut = _serve(reqs, False, batch_cutoff=-1, max_batch=16)
    assert cut >= 32
    for r in reqs:
        assert comps[r["id"]]["status"] == 0
        ref = np.linalg.solve(r["a"], r["b"])
        assert np.linalg.norm(xs[r["id"]] - ref) <= 1e-10 * np.linalg.norm(ref)


def test_server_create_destroy_returns_device_memory():
    """Five servers created, used and destroyed: the workers' solve scratch goes with them, so
    free device memory returns to where it was after the first (which warms the per-device
    factorization workspaces that outlive servers by design)."""
    torch = _torch()
    reqs = _requests(7, 12, [4096, 3000, 64, 2048])

    def cycle():
        _serve(reqs, False, batch_cutoff=128)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        return torch.cuda.mem_get_info()[0]

    cycle()
    base = cycle()
    free = [cycle() for _ in range(4)]
    print("free bytes after each cycle:", base, free, flush=True)
    assert min(free) >= base - (16 << 20)


@pytest.mark.parametrize("mixed", [False, True])
def test_server_graph_replayed_sweeps_equal_eager_sweeps(mixed):
    """Later right-hand sides of a resident factor replay the factor's refinement sweep from a
    HIP graph recorded at its first reuse.  Against the same requests with every sweep launched
    kernel by kernel: every solution bitwise equal and the same sweep counts.  Each factor's
    graph is replayed by several requests with other work between the replays (fresh
    factorizations, batched rounds, the other factors' graphs); a request with another nrhs
    records the graph again; run once request by request (every reuse replays the graph) and
    once all at once (concurrent reuses of one factor take turns or sweep eagerly)."""
    rng = np.random.default_rng(91)
    sizes = [300, 1000, 2100]
    reqs, fid = [], {}

    def add(n, nrhs, reuse, keep, a=None):
        i = len(reqs)
        if a is None:
            a = reqs[reuse]["a"] if reuse >= 0 else workloads.normal(n, n, 4000 + i)
        reqs.append(dict(id=i, n=n, nrhs=nrhs, a=a, reuse=reuse, keep=keep,
                         b=np.asfortranarray(rng.standard_normal((n, nrhs)))))
        return i

    for n in sizes:
        fid[n] = add(n, 2, -1, True)
    for rnd in range(4):
        for n in sizes:
            add(n, 5 if rnd == 2 else 2, fid[n], False)
        add(500 + 10 * rnd, 3, -1, False)  # a fresh factorization between the replays
        add(64, 2, -1, False)              # and a batched round
    ce, xe, _ = _serve(reqs, True, batch_cutoff=128, mixed=mixed, graphs=False)
    cg, xg, _ = _serve(reqs, True, batch_cutoff=128, mixed=mixed, graphs=True)
    cc, xc, _ = _serve(reqs, False, batch_cutoff=128, mixed=mixed, graphs=True)
    replays = 0
    for r in reqs:
        i = r["id"]
        for c in (ce[i], cg[i], cc[i]):
            assert c["status"] == 0, (i, c)
        assert np.array_equal(xe[i], xg[i]), f"request {i} (n={r['n']}): graph != eager"
        assert np.array_equal(xe[i], xc[i]), f"request {i} (n={r['n']}): concurrent != eager"
        assert ce[i]["iterations"] == cg[i]["iterations"] == cc[i]["iterations"], i
        assert ce[i]["graph_sweeps"] == 0
        if r["reuse"] >= 0:
            assert cg[i]["reused"] == 1
            assert cg[i]["graph_sweeps"] == cg[i]["iterations"] >= 1, (i, cg[i])
            replays += 1
        else:
            assert cg[i]["graph_sweeps"] == 0
        x = xe[i]
        res = np.abs(r["b"] - r["a"] @ x).max(axis=0)
        den = np.abs(r["a"]).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(r["b"]).max(axis=0)
        assert (res / den).max() <= 1e-13, i
    assert replays == 4 * len(sizes)
    print(f"mixed={mixed}: {replays} graph-replayed requests, sweeps per request "
          f"{sorted({cg[r['id']]['iterations'] for r in reqs if r['reuse'] >= 0})}", flush=True)
