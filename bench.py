#!/usr/bin/env python3
"""Benchmark of the plain FP64 delayed LU on one MI355X (device-memory interface).

Workload (one "step" = one factorization): an n x n column-major FP64 matrix, entries iid N(0,1)
from a fixed seed, already resident in HBM; before each step the pristine copy is restored with
a device-to-device copy (outside the factorization's own timing, as in the reference's
delayed_lu_cuda/benchmark.cu:45-57).  value = (2/3) n^3 / mean factorization time, in TFLOP/s.

Also reported on the same JSON line:
  roofline      for the dominant kernel (the FP64 MFMA trailing update): algorithmic flops per
                launch / mean launch duration, both measured live with HIP events on the stream
                each launch runs on (one profiled factorization after the timed steps);
  cpu_baseline  the oracle (C restatement of the reference delayed_getrf on the host BLAS) on the
                same n = 16384 matrix, timed on the host threads one GPU gets, plus LAPACK
                dgetrf on it;
  parity        permutations (bit for bit) and factors (relative Frobenius difference) of the
                timed factorization versus the oracle's.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

# The solve service runs three full-size workers, each factorization on two internal streams
# beside the worker's own: they overlap only on streams with hardware queues of their own (HIP's
# default is 4 per process; measured on the serving run: 516 requests/s closed loop with 4,
# 636-652 with 8, no gain from 16).  Read at HIP initialisation, so set before torch is imported;
# an explicit setting in the environment wins.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "delayed-pivoting_amd")
FP64_PEAK_TFLOPS = 78.6


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_and_parity(a, lu_gpu, first, threads):
    """The oracle (C restatement of the reference delayed_getrf, nb = 64, on the host BLAS) and
    LAPACK dgetrf on the full benchmark matrix, each timed once on `threads` host threads: the
    CPU baseline, and the parity check of the timed factorization (permutations bit for bit,
    factors within 1e-12 relative)."""
    import numpy as np
    oracle = _load("dlu_oracle_py", os.path.join(ROOT, "oracle", "oracle.py"))
    oracle.set_threads(threads)
    n = a.shape[0]
    flops = 2.0 * n ** 3 / 3.0
    t0 = time.perf_counter()
    lu_c, ip_c, jp_c, info_c, ex = oracle.delayed_getrf(a, nb=64, margins=True)
    t_delayed = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, ipl, infol = oracle.lapack_getrf(a)
    t_lapack = time.perf_counter() - t0
    ipiv, jpiv, rep = first
    same = bool(np.array_equal(ipiv, ip_c) and np.array_equal(jpiv, jp_c) and info_c == n)
    parity = {"n": n, "bit_exact_permutations": same,
              "factor_rel_diff_fro": float(np.linalg.norm(lu_gpu - lu_c) / np.linalg.norm(lu_c)),
              "cpu_min_margin": float(ex["margins"].min()), "gpu_min_margin": rep["min_margin"],
              "gpu_near_ties": rep["near_ties"], "gpu_exact_ties": rep["exact_ties"]}
    if not same:
        k = int(np.argmax((ipiv != ip_c) | (jpiv != jp_c)))
        parity["first_differing_step"] = k
        parity["cpu_margin_at_first_difference"] = float(ex["margins"][k])
    cpu = {
        "value": round(flops / t_delayed / 1e12, 4), "unit": "TFLOP/s", "cores": threads,
        "kind": "port",
        "sample": f"the whole workload: one delayed_getrf restatement (nb=64) on the same "
                  f"{n}x{n} N(0,1) matrix, {oracle.blas_name().strip()[:60]}, {threads} threads "
                  f"(the CPUs this host gives one GPU)",
        "seconds": round(t_delayed, 3), "info": info_c,
        "lapack_dgetrf": {"value": round(flops / t_lapack / 1e12, 4), "unit": "TFLOP/s",
                          "seconds": round(t_lapack, 3),
                          "same_ipiv": bool(np.array_equal(ip_c, ipl))},
    }
    return cpu, parity


def _timed(torch, fn, reps=2):
    """Best of `reps` after one warm-up, host wall clock around a synchronized call."""
    best = None
    for i in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = fn()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        if i > 0 and (best is None or t < best[0]):
            best = (t, res)
    return best


def bench_delay_stress(dlu, torch, dev, n, column=False):
    """Delay-stress FP64: 5% zero columns, 5% scaled by 1e-14, a rank-32 64x64 diagonal block;
    the full-pivot fallback (delayed_getrf) or the column-delay variant (delayed_getrf_column)."""
    import numpy as np
    workloads = _load("dlu_workloads", os.path.join(PKG, "workloads.py"))
    log(f"delay-stress n={n} column={column}")
    a, zero_cols, _, _ = workloads.delay_stress(n, n)
    orig = torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)
    work = torch.empty_like(orig)
    fb = dlu.FALLBACK_COLUMN if column else dlu.FALLBACK_FULL

    def run():
        work.copy_(orig)
        return dlu.getrf_device(n, n, work.data_ptr(), n, options={"nb": 0, "fallback": fb})

    t, (st, ipiv, jpiv, info, rep, _) = _timed(torch, run)
    delayed = np.sort(dlu.net_order(jpiv, info, n)[info:])
    # One more, profiled run: the full-pivot searches' own time and the bytes they must read.
    work.copy_(orig)
    _, _, _, _, prep, _ = dlu.getrf_device(n, n, work.data_ptr(), n,
                                           options={"nb": 0, "profile": 1, "fallback": fb})
    torch.cuda.synchronize()
    search = {"launches": prep["search_launches"], "ms": round(prep["search_seconds"] * 1e3, 2),
              "materialize_and_update_ms": round(prep["update_seconds"] * 1e3, 2),
              "panel_ms": round(prep["panel_seconds"] * 1e3, 2)}
    if column:
        # The column search stops at the first column (from the right) with a nonzero in the
        # active rows, so the bytes it reads depend on the input; no rate is claimed for it.
        search["reads"] = "right to left until a column holds a nonzero in the active rows"
    else:
        search["bytes"] = prep["search_bytes"]
        search["gb_per_s"] = round(prep["search_bytes"] / max(prep["search_seconds"], 1e-12) / 1e9, 1)
    variant = "delayed_getrf_column (rightmost nonzero column)" if column else \
        "delayed_getrf (full-pivot search)"
    return {"workload": f"delay-stress FP64 n={n} (5% zero cols, 5% x1e-14, rank-32 64x64 block), "
                        f"{variant}",
            "ms": round(t * 1e3, 2), "tflops_2n3_3": round(2.0 * n ** 3 / 3 / t / 1e12, 3),
            "info": info, "expected_info": n - len(zero_cols),
            "delayed_set_equals_zero_columns": bool(info == n - len(zero_cols) and
                                                    np.array_equal(delayed, zero_cols)),
            "fallbacks": rep["fallbacks"], "panels": rep["panels"], "status": int(st),
            ("column_search" if column else "full_pivot_search"): search,
            "time_includes": "device copy of the input (ms-scale) + factorization"}


def bench_mixed(dlu, torch, dev, n, nrhs):
    """FP32 factor + FP64 refinement vs plain FP64 on A = U diag(s) V^T, s from 1 to 1e-6."""
    log(f"mixed precision n={n}, {nrhs} RHS")
    g = torch.Generator(device=dev)
    g.manual_seed(n)
    u = torch.linalg.qr(torch.randn((n, n), dtype=torch.float64, device=dev, generator=g))[0]
    v = torch.linalg.qr(torch.randn((n, n), dtype=torch.float64, device=dev, generator=g))[0]
    s = torch.tensor(1e-6, dtype=torch.float64, device=dev) ** (
        torch.arange(n, dtype=torch.float64, device=dev) / (n - 1))
    at = ((u * s) @ v.T).T.contiguous()  # storage rows = columns of A
    del u, v
    b = torch.randn((nrhs, n), dtype=torch.float64, device=dev, generator=g)
    x = torch.zeros_like(b)
    out = {"workload": f"A = U diag(s) V^T, n={n}, cond 1e6, {nrhs} RHS N(0,1)", "tol": 1e-13}
    for mixed in (False, True):
        t, (st, rep, bwd) = _timed(torch, lambda: dlu.gesv_device(
            n, nrhs, at.data_ptr(), n, b.data_ptr(), n, x.data_ptr(), n, mixed=mixed))
        key = "mixed" if mixed else "fp64"
        out[key] = {"ms": round(t * 1e3, 2), "factor_ms": round(rep["factor_seconds"] * 1e3, 2),
                    "refine_ms": round(rep["refine_seconds"] * 1e3, 2),
                    "iterations": rep["iterations"], "converged": bool(rep["converged"]),
                    "max_backward_error": float(bwd.max()), "status": int(st)}
    out["speedup_mixed_over_fp64"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    del at, b, x
    torch.cuda.empty_cache()
    return out


def bench_sytrf(dlu, torch, dev, sizes, cpu_n, threads):
    """The delayed Bunch-Kaufman LDL^T (delayed_sytrf_blocked) on symmetric N(0,1) matrices
    A = X + X^T: GPU time per n (best of two after a warm-up), n^3/3 flops counted; the oracle
    (C restatement, MKL/OpenBLAS on the host) on the cpu_n matrix, with the permutations and
    block sizes compared."""
    import numpy as np
    out = {"workload": "delayed Bunch-Kaufman LDL^T, A = X + X^T, X N(0,1) seeded, device "
                       "interface, panels of 64", "flops_counted": "n^3/3", "gpu": []}
    for n in sizes:
        log(f"sytrf n={n}")
        g = torch.Generator(device=dev)
        g.manual_seed(n + 7)
        x = torch.randn((n, n), dtype=torch.float64, device=dev, generator=g)
        orig = (x + x.T).contiguous()  # symmetric: storage rows = columns
        del x
        work = torch.empty_like(orig)

        def run():
            work.copy_(orig)
            return dlu.sytrf_device(n, work.data_ptr(), n)

        t, (st, perm, piv, info) = _timed(torch, run)
        out["gpu"].append({"n": n, "ms": round(t * 1e3, 2),
                           "gflops": round(n ** 3 / 3.0 / t / 1e9, 1), "info": info,
                           "two_by_two": int((piv == 2).sum()), "status": int(st)})
        if n == cpu_n and threads > 0:
            oracle = _load("dlu_oracle_py", os.path.join(ROOT, "oracle", "oracle.py"))
            oracle.set_threads(threads)
            a = np.asfortranarray(orig.cpu().numpy())
            t0 = time.perf_counter()
            _, pc, qc, ic = oracle.delayed_sytrf(a)
            tc = time.perf_counter() - t0
            out["cpu_oracle"] = {"n": n, "ms": round(tc * 1e3, 1), "threads": threads,
                                 "gflops": round(n ** 3 / 3.0 / tc / 1e9, 1),
                                 "same_perm_and_blocks": bool(np.array_equal(pc, perm) and
                                                              np.array_equal(qc, piv) and
                                                              ic == info)}
        del orig, work
        torch.cuda.empty_cache()
    return out


def bench_sweep(dlu, torch, dev, sizes):
    """Plain FP64 delayed LU over the size sweep of the headline configuration: best of two
    factorizations after one warm-up, N(0,1) seeded per n, device interface."""
    out = []
    for n in sizes:
        log(f"sweep n={n}")
        g = torch.Generator(device=dev)
        g.manual_seed(n)
        orig = torch.randn((n, n), dtype=torch.float64, device=dev, generator=g)
        work = torch.empty_like(orig)
        best = None
        info = None
        for rep in range(3):
            work.copy_(orig)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st, _, _, info, _, _ = dlu.getrf_device(n, n, work.data_ptr(), n)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if rep > 0 and int(st) == 0:
                best = t if best is None else min(best, t)
        out.append({"n": n, "ms": round(best * 1e3, 3) if best else None,
                    "tflops": round(2.0 * n ** 3 / 3.0 / best / 1e12, 3) if best else None,
                    "info": info})
        del orig, work
        torch.cuda.empty_cache()
    return out


def bench_serving(dlu, torch, dev, count, seed=20000):
    """Poisson arrivals of factor + solve requests, n log-uniform in [64, 4096], 1-8 RHS,
    30% reusing an earlier request's resident factor; FP64 and mixed-precision services.
    Each new request's matrix is its own window (leading dimension 4096) of one N(0,1) pool of
    4096 x 32768 doubles at a random row and column offset: the service only reads A, and
    20000 separate matrices would need about 226 GB."""
    import math
    import numpy as np
    rng = np.random.default_rng(seed)
    ld, pool_cols = 4096, 32768
    reqs = []
    for i in range(count):
        if i >= 8 and rng.random() < 0.3:
            j = int(rng.integers(0, i))
            while reqs[j]["reuse"] >= 0:
                j = reqs[j]["reuse"]
            reqs.append({"n": reqs[j]["n"], "reuse": j, "nrhs": int(rng.integers(1, 9))})
        else:
            n = int(round(math.exp(rng.uniform(math.log(64), math.log(4096)))))
            reqs.append({"n": n, "reuse": -1, "nrhs": int(rng.integers(1, 9)),
                         "off": int(rng.integers(0, ld - n + 1))
                         + ld * int(rng.integers(0, pool_cols - n + 1))})
    log(f"serving: {count} requests, generating inputs on the device")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    pool = torch.randn((pool_cols, ld), dtype=torch.float64, device=dev, generator=g)
    for r in reqs:
        r["b"] = torch.randn((r["nrhs"], r["n"]), dtype=torch.float64, device=dev, generator=g)
        r["x"] = torch.zeros_like(r["b"])
    torch.cuda.synchronize()
    flops = sum((0 if r["reuse"] >= 0 else 2.0 * r["n"] ** 3 / 3) + 2.0 * r["n"] ** 2 * r["nrhs"]
                for r in reqs)

    def a_of(i):
        """Device address of request i's matrix (leading dimension ld)."""
        owner = reqs[i]["reuse"] if reqs[i]["reuse"] >= 0 else i
        return pool.data_ptr() + 8 * reqs[owner]["off"]

    def run(srv, arrivals, upto=count):
        t0 = srv.now()
        for i, r in enumerate(reqs[:upto]):
            if arrivals is not None:
                while srv.now() - t0 < arrivals[i]:
                    time.sleep(min(0.0005, max(0.0, arrivals[i] - (srv.now() - t0))))
            n = r["n"]
            st = srv.submit(i, n, r["nrhs"], a_of(i), ld, r["b"].data_ptr(), n,
                            r["x"].data_ptr(), n, factor_of=r["reuse"], keep=r["reuse"] < 0)
            if st != dlu.Status.SUCCESS:
                raise RuntimeError(f"submit {i}: {dlu.status_string(st)}")
        srv.drain()
        comps = {c["id"]: c for c in srv.poll(upto + 16)}
        return t0, comps

    out = {"workload": f"{count} requests" + ("" if count == 20000 else
                                              f" (the first {count} of the seeded stream; the "
                                              "full configuration has 20000)")
                       + ", n log-uniform [64,4096], 1-8 RHS, 30% reuse; matrices are windows "
                         "(lda 4096) of one N(0,1) pool",
           "flops": flops}
    srv = dlu.Server(budget_gib=64.0, mixed=False)
    cutoff = srv.cutoff
    # capacity of the FP64 service, closed loop over the stream's first 5000 requests
    probe = min(count, 5000)
    t0, comps = run(srv, None, probe)
    span = max(c["t_done"] for c in comps.values()) - t0
    capacity = probe / span
    rate = 0.7 * capacity
    srv.close()
    out.update({"batch_cutoff": cutoff, "closed_loop_requests_per_s": round(capacity, 1),
                "arrival_rate_per_s": round(rate, 1)})
    arrivals = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, count))
    for mixed in (False, True):
        for r in reqs:
            r["x"].zero_()
        torch.cuda.synchronize()
        srv = dlu.Server(budget_gib=64.0, mixed=mixed, batch_cutoff=cutoff)
        t0, comps = run(srv, arrivals)
        srv.close()
        lat = np.array([comps[i]["t_done"] - (t0 + arrivals[i]) for i in range(count)])
        ok = all(comps[i]["status"] == 0 for i in range(count))
        worst = 0.0
        for i in range(0, count, max(1, count // 200)):  # a sample, checked on the device
            n = reqs[i]["n"]
            st, bw = dlu.backward_error_device(n, reqs[i]["nrhs"], a_of(i), ld,
                                               reqs[i]["x"].data_ptr(), n,
                                               reqs[i]["b"].data_ptr(), n)
            worst = max(worst, float(bw.max()))
        span = max(c["t_done"] for c in comps.values()) - t0
        out["mixed" if mixed else "fp64"] = {
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "mean_ms": round(float(lat.mean()) * 1e3, 3),
            "solves_per_s": round(count / span, 1),
            "tflops": round(flops / span / 1e12, 3),
            "batched": sum(1 for c in comps.values() if c["path"] == 0),
            "reused": sum(1 for c in comps.values() if c["reused"]),
            "graph_replayed": sum(1 for c in comps.values() if c["graph_sweeps"] > 0),
            "all_succeeded": ok, "max_backward_error_sampled": worst}
        log(f"serving mixed={mixed}: p50 {out['mixed' if mixed else 'fp64']['p50_ms']} ms "
            f"p99 {out['mixed' if mixed else 'fp64']['p99_ms']} ms")
    del pool, reqs
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--nb", type=int, default=0, help="panel width (0: library default)")
    ap.add_argument("--seed", type=int, default=16384)
    ap.add_argument("--panel-cus", type=int, default=0,
                    help="CUs reserved for panel kernels (0 auto, -1 none)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0: OMP_NUM_THREADS, else up to 16)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline and the parity check against the oracle")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the delay-stress, mixed-precision and serving measurements")
    ap.add_argument("--extra", default="sweep,c3,c4,c5,sytrf",
                    help="which extra measurements to run (comma-separated of sweep, c3, c4, "
                         "c5, sytrf), e.g. c4 alone for a profiled run of the mixed-precision "
                         "solve")
    ap.add_argument("--mixed-n", type=int, default=16384)
    ap.add_argument("--serve-requests", type=int, default=20000,
                    help="requests of the serving stream measured per precision mode")
    ap.add_argument("--gemm", type=int, nargs=2, metavar=("M", "K"), default=None,
                    help="only time the trailing-update kernel C(M x M) -= A(M x K) B(K x M) "
                         "(the library's dlu_gemm_sub_device), for counter runs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device to device three times (the reference kernel "
                         "that scales FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the factorization runs on a single MI355X).", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    if args.calibrate_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": (1 << 27) * 8, "copies": 3}), flush=True)
        return 0
    dlu = _load("dlu_amd", os.path.join(PKG, "dlu.py"))
    dlu.lib()
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.gemm:
        # C, A and B column-major with leading dimension M (rows of each tensor are columns).
        m, kk = args.gemm
        c = torch.randn((m, m), dtype=torch.float64, device="cuda")
        a = torch.randn((kk, m), dtype=torch.float64, device="cuda")
        b = torch.randn((m, m), dtype=torch.float64, device="cuda")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        best = float("inf")
        for _ in range(max(1, args.steps)):
            e0.record()
            st = dlu.gemm_sub_device(m, m, kk, c.data_ptr(), a.data_ptr(), b.data_ptr(), m)
            e1.record()
            torch.cuda.synchronize()
            if st != dlu.Status.SUCCESS:
                raise RuntimeError(dlu.status_string(st))
            best = min(best, e0.elapsed_time(e1) / 1e3)
        print(json.dumps({"gemm": {"M": m, "N": m, "K": kk, "best_ms": round(best * 1e3, 4),
                                   "tflops": round(2.0 * m * m * kk / best / 1e12, 2)}}), flush=True)
        return 0
    dev = torch.device("cuda", 0)
    n, nb = args.n, args.nb
    flops = 2.0 * n ** 3 / 3.0

    log(f"generating {n}x{n} N(0,1) on the device (seed {args.seed})")
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    orig = torch.randn((n, n), dtype=torch.float64, device=dev, generator=g)  # row j = column j
    work = torch.empty_like(orig)
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)

    def step(profile=False):
        work.copy_(orig)
        ev0.record(stream)
        # The profiled (untimed) run also tracks the per-step margins for the tie report.
        st, ipiv, jpiv, info, rep, _ = dlu.getrf_device(
            n, n, work.data_ptr(), n, stream=stream,
            options={"nb": nb, "profile": 1 if profile else 0, "panel_cus": args.panel_cus},
            margins=profile)
        ev1.record(stream)
        torch.cuda.synchronize()
        if st != dlu.Status.SUCCESS or info != n:
            raise RuntimeError(f"factorization failed: {dlu.status_string(st)} info={info}")
        return ev0.elapsed_time(ev1) / 1e3, ipiv, jpiv, rep

    for i in range(args.warmup):
        t, *_ = step()
        log(f"warmup {i}: {t * 1e3:.2f} ms")
    torch.cuda.synchronize()
    times = []
    first = None
    wall0 = time.perf_counter()
    for i in range(args.steps):
        t, ipiv, jpiv, rep = step()
        times.append(t)
        if first is None:
            first = (ipiv, jpiv, rep)
        log(f"step {i}: {t * 1e3:.2f} ms  {flops / t / 1e12:.2f} TFLOP/s")
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    mean_t = sum(times) / len(times)
    value = flops / mean_t / 1e12

    # One profiled factorization: per-launch HIP events on each kernel's own stream.
    _, pipiv, pjpiv, prep = step(profile=True)
    same_run = bool(np.array_equal(pipiv, first[0]) and np.array_equal(pjpiv, first[1]))
    first = (first[0], first[1], prep)  # its tie report (the factorization is deterministic)
    torch.cuda.synchronize()
    # dominant kernel: the trailing-update GEMM (launches over at least 512 128 x 128 blocks of C)
    upd_launch_s = prep["large_update_seconds"] / max(1, prep["large_update_launches"])
    upd_flops = prep["large_update_flops"] / max(1, prep["large_update_launches"])
    achieved = upd_flops / upd_launch_s / 1e12 if upd_launch_s > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("gemm_sub_kernel_bytes_per_launch")
    alg_bytes = prep.get("large_update_bytes", 0.0) / max(1, prep["large_update_launches"])
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                "traffic": traffic, "algorithmic_bytes_per_launch": round(alg_bytes),
                "kernel": "gemm_sub_kernel<double,128,128,16,true,2,false> (A22 -= L21*U12, FP64 MFMA, "
                          "K = 128 single panels and K = 256 panel pairs)",
                "launches": prep["large_update_launches"],
                "mean_launch_ms": round(upd_launch_s * 1e3, 4),
                "flops_per_launch": upd_flops}
    phases = {"panel_ms_total": round(prep["panel_seconds"] * 1e3, 3),
              "panel_launches": prep["panel_launches"],
              "solve_ms_total": round(prep["solve_seconds"] * 1e3, 3),
              "update_ms_total": round(prep["update_seconds"] * 1e3, 3)}

    parity = cpu = None
    if not args.no_cpu:
        # The last (profiled) step left the factors in `work`; the factorization is deterministic,
        # so they are those of every timed step.
        a = np.asfortranarray(orig.cpu().numpy().T)  # column-major view of the same matrix
        lu_gpu = np.asfortranarray(work.cpu().numpy().T)
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
            min(16, os.cpu_count() or 1)
        log(f"cpu baseline and parity: oracle and LAPACK dgetrf on the {n}^2 matrix, "
            f"{threads} threads")
        cpu, parity = cpu_and_parity(a, lu_gpu, first, threads)
        del a, lu_gpu

    del orig, work
    torch.cuda.empty_cache()
    c3 = c4 = c5 = sweep = sytrf = None
    extra = set() if args.no_extra else {e.strip() for e in args.extra.split(",") if e.strip()}
    if "sweep" in extra:
        sweep = bench_sweep(dlu, torch, dev, [2048, 4096, 8192, 16384, 32768])
    if "c3" in extra:
        c3 = bench_delay_stress(dlu, torch, dev, 8192)
        c3["column_variant"] = bench_delay_stress(dlu, torch, dev, 8192, column=True)
    if "c4" in extra:
        c4 = bench_mixed(dlu, torch, dev, args.mixed_n, 16)
    if "c5" in extra:
        c5 = bench_serving(dlu, torch, dev, args.serve_requests)
    if "sytrf" in extra:
        sytrf = bench_sytrf(dlu, torch, dev, [4096, 8192], 4096,
                            0 if args.no_cpu else (args.cpu_threads or int(
                                os.environ.get("OMP_NUM_THREADS", "0") or 0) or
                                min(16, os.cpu_count() or 1)))

    out = {
        "metric": "FP64 LU TFLOP/s @n=16384 (% peak); FP32+IR speedup @bwd err<=1e-13; p99 ms",
        "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(mean_t * 1e3, 3),
        "ms_per_step_wall": round(wall / args.steps * 1e3, 3),
        "pct_fp64_peak": round(100.0 * value / FP64_PEAK_TFLOPS, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic N(0,1), seeded, resident in HBM",
        "config": {"workload": f"plain FP64 delayed LU, n={n}, nb={nb or 128}, device interface",
                   "n": n, "nb": nb or 128, "flops_counted": "2n^3/3"},
        "roofline": roofline, "phases": phases, "cpu_baseline": cpu, "parity": parity,
        "pivot_report": dict({k: first[2][k] for k in ("fallbacks", "near_ties", "exact_ties",
                                                       "min_margin", "panels")},
                             source="the profiled run, margins tracked; its permutations equal "
                                    "the timed runs'" if same_run else
                                    "the profiled run (its permutations differ from the timed "
                                    "runs')") if first[2] else None,
        "c2_sweep": sweep, "c3_delay_stress": c3, "c4_mixed_precision": c4, "c5_serving": c5,
        "f4_sytrf": sytrf,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
