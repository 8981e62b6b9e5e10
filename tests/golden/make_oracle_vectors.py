"""Write tests/golden/oracle_vectors.json: pivots of seeded inputs computed by the oracle.

The reference's C++ sources cannot be compiled in this image (no cblas.h), so these vectors come
from the oracle restatement (oracle/dlu_oracle.c on MKL).  They are pinned independently by the
checks in tests/test_oracle.py (LAPACK dgetrf agreement, long-double brute force, closed forms)
and make any later drift of the oracle or of the host BLAS visible.
Run: python tests/golden/make_oracle_vectors.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import oracle, sha256, workloads  # noqa: E402

out = {"blas": oracle.blas_name().strip(), "cases": []}

a = workloads.normal(512, 512, 1)  # C1: 512 x 512, N(0,1), seed 1, nb = 64
lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a, nb=64, margins=True)
out["cases"].append({
    "name": "C1 normal 512 seed 1", "kind": "normal", "m": 512, "n": 512, "seed": 1,
    "sha256": sha256(a), "info": info, "ipiv": ipiv.tolist(), "jpiv": jpiv.tolist(),
    "fallbacks": ex["fallbacks"], "min_margin": float(ex["margins"].min()),
    "backward_error": oracle.backward_error(a, lu, ipiv, jpiv, info),
    "lu_checksum": float(lu.sum())})

for n, seed in ((512, 3), (1024, 5)):
    a, zc, tc, off = workloads.delay_stress(n, seed)
    lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a, nb=64)
    out["cases"].append({
        "name": f"delay-stress {n} seed {seed}", "kind": "delay_stress", "m": n, "n": n,
        "seed": seed, "sha256": sha256(a), "info": info, "ipiv": ipiv.tolist(),
        "jpiv": jpiv.tolist(), "fallbacks": ex["fallbacks"], "zero_cols": zc.tolist(),
        "backward_error": oracle.backward_error(a, lu, ipiv, jpiv, info)})

for (m, n, seed) in ((300, 200, 7), (200, 300, 8)):
    a = workloads.normal(m, n, seed)
    lu, ipiv, jpiv, info, ex = oracle.delayed_getrf(a, nb=64)
    out["cases"].append({
        "name": f"normal {m}x{n} seed {seed}", "kind": "normal", "m": m, "n": n, "seed": seed,
        "sha256": sha256(a), "info": info, "ipiv": ipiv.tolist(), "jpiv": jpiv.tolist(),
        "fallbacks": ex["fallbacks"],
        "backward_error": oracle.backward_error(a, lu, ipiv, jpiv, info)})

path = os.path.join(HERE, "oracle_vectors.json")
with open(path, "w") as f:
    json.dump(out, f)
print(path)
