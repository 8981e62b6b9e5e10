"""Autotuner of libdlu: times the trailing-update GEMM's tile configurations and the batched
kernel against the blocked path on the current GPU, and writes the table the library compiles in
(csrc/dlu_tuning.inc).  Rebuild the library afterwards.

    python delayed-pivoting_amd/autotune.py [--out PATH] [--json PATH]

GEMM: every tile configuration (64 x 64, 128 x 128, 128 x 64, 64 x 128) is timed, best of three
after a warm-up, on shapes that the factorization issues in each size class (counted in 128 x 128
blocks of C) at the two update depths (K = 128 single panels, K = 256 panel pairs); a class takes
the configuration with the least summed time.  Batch cutoff: 128 matrices of order n through the
one-workgroup-per-matrix kernel against the same matrices one at a time through the blocked
factorization and solve; the cutoff is the largest n where the batched kernel costs less per
matrix.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIG_NAMES = {0: "64x64", 1: "128x128", 2: "128x64", 3: "64x128"}
# Representative (M, N) of each size class, from the factorization's own launches: the next
# panel's block (N = nb) and the far update (square-ish), rows of the trailing matrix.
SIZE_CLASSES = [
    ("< 64 blocks", [(1024, 128), (896, 896)]),
    ("< 256 blocks", [(16000, 128), (1920, 1920)]),
    ("< 512 blocks", [(8192, 640), (2816, 2816)]),
    ("< 2048 blocks", [(16000, 1536), (5632, 5632)]),
    (">= 2048 blocks", [(16128, 8064), (12000, 12000)]),
]
DEPTHS = [128, 256]


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def time_gemm(dlu, torch, M, N, K, config, reps=3):
    c = torch.randn((N, M), dtype=torch.float64, device="cuda")   # column-major, ld = M
    a = torch.randn((K, M), dtype=torch.float64, device="cuda")
    b = torch.randn((N, M), dtype=torch.float64, device="cuda")   # K x N with ld = M
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for i in range(reps + 1):
        e0.record()
        st = dlu.gemm_sub_device(M, N, K, c.data_ptr(), a.data_ptr(), b.data_ptr(), M,
                                 config=config)
        e1.record()
        torch.cuda.synchronize()
        if st != dlu.Status.SUCCESS:
            raise RuntimeError(f"gemm config {config}: {dlu.status_string(st)}")
        if i > 0:
            best = min(best, e0.elapsed_time(e1) * 1e-3)
    del a, b, c
    return best


def tune_gemm(dlu, torch, log):
    table, detail = [], {}
    for K in DEPTHS:
        row = []
        for label, shapes in SIZE_CLASSES:
            totals = {}
            for cfg in CONFIG_NAMES:
                totals[cfg] = sum(time_gemm(dlu, torch, M, N, K, cfg) for M, N in shapes)
            best = min(totals, key=totals.get)
            row.append(best)
            detail[f"K={K} {label}"] = {CONFIG_NAMES[c]: round(t * 1e3, 4) for c, t in totals.items()}
            log(f"K={K} {label}: best {CONFIG_NAMES[best]}  "
                + "  ".join(f"{CONFIG_NAMES[c]} {t * 1e3:.3f} ms" for c, t in totals.items()))
        table.append(row)
    torch.cuda.empty_cache()
    return table, detail


def tune_cutoff(dlu, torch, log, count=128):
    best_n, detail = 0, {}
    for n in (32, 48, 64, 80, 96, 112, 128):
        a = torch.randn((count, n, n), dtype=torch.float64, device="cuda")
        b = torch.randn((count, 1, n), dtype=torch.float64, device="cuda")
        lu = torch.empty_like(a)
        x = torch.empty_like(b)
        piv = torch.zeros((count, 2 * n), dtype=torch.int32, device="cuda")
        info = torch.zeros(count, dtype=torch.int32, device="cuda")
        items = [dlu.BatchItem(a=a[i].data_ptr(), lda=n, lu=lu[i].data_ptr(), ldlu=n,
                               ipiv=piv[i].data_ptr(), jpiv=piv[i].data_ptr() + 4 * n,
                               info=info[i:].data_ptr(), info_in=0, n=n, nrhs=1,
                               b=b[i].data_ptr(), ldb=n, x=x[i].data_ptr(), ldx=n, solve_only=0)
                 for i in range(count)]
        torch.cuda.synchronize()
        tb = []
        for _ in range(3):
            t0 = time.perf_counter()
            st = dlu.batch_gesv_device(items)
            tb.append(time.perf_counter() - t0)
            if st != dlu.Status.SUCCESS:
                raise RuntimeError(dlu.status_string(st))
        tl = []
        for i in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st, _, _ = dlu.gesv_device(n, 1, a[i].data_ptr(), n, b[i].data_ptr(), n,
                                       x[i].data_ptr(), n, mixed=False)
            tl.append(time.perf_counter() - t0)
        per_batched, per_full = min(tb) / count, min(tl)
        detail[n] = {"batched_us_per_matrix": round(per_batched * 1e6, 2),
                     "blocked_us": round(per_full * 1e6, 2)}
        log(f"n={n}: batched {per_batched * 1e6:.1f} us per matrix, blocked {per_full * 1e6:.1f} us")
        if per_batched <= per_full:
            best_n = n
        del a, b, lu, x, piv, info
    torch.cuda.empty_cache()
    return max(best_n, 32), detail


def render(table, cutoff, source):
    rows = ", ".join("{" + ", ".join(str(c) for c in row) + "}" for row in table)
    return f"""// Tuned parameters of libdlu, written by delayed-pivoting_amd/autotune.py from timings on the
// device named below (rerun it on other hardware and rebuild).
//
// kTunedGemmTiles[depth][size]: the trailing-update GEMM's tile configuration
//   depth: 0 for K <= 160 (single panels), 1 for deeper updates (panel pairs, K = 256)
//   size:  the launch's count of 128 x 128 blocks of C: < 64, < 256, < 512, < 2048, >= 2048
//   configuration: 0 = 64 x 64 tiles (4 workgroups per CU), 1 = 128 x 128 (2 per CU),
//                  2 = 128 x 64 (3 per CU), 3 = 64 x 128 (3 per CU)
// kTunedBatchCutoff: the largest n the solve service sends to the one-workgroup-per-matrix
//   batched kernel.
#pragma once
namespace dlu {{
constexpr int kTunedGemmTiles[2][5] = {{{rows}}};
constexpr int kTunedBatchCutoff = {cutoff};
constexpr const char *kTunedSource = "{source}";
}}  // namespace dlu
"""


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "dlu_tuning.inc"))
    ap.add_argument("--json", default=None, help="also write the timings here")
    args = ap.parse_args()
    import torch
    assert torch.cuda.is_available(), "the autotuner needs a GPU"
    dlu = _load("dlu_amd", os.path.join(HERE, "dlu.py"))
    dlu.lib()

    def log(msg):
        print(msg, flush=True)

    table, gemm_detail = tune_gemm(dlu, torch, log)
    cutoff, cut_detail = tune_cutoff(dlu, torch, log)
    source = f"{torch.cuda.get_device_name(0)}, {time.strftime('%Y-%m-%d')}"
    with open(args.out, "w") as f:
        f.write(render(table, cutoff, source))
    log(f"wrote {args.out}: tiles {table}, batch cutoff {cutoff}")
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"device": source, "gemm_ms": gemm_detail, "batch_cutoff": cut_detail,
                       "table": table, "cutoff": cutoff}, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
