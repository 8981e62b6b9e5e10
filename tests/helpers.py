"""Shared test helpers: load the oracle, the libdlu binding and the workload generators."""
import hashlib
import importlib.util
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "delayed-pivoting_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


oracle = _load("dlu_oracle_py", os.path.join(ROOT, "oracle", "oracle.py"))
dlu = _load("dlu_amd", os.path.join(PKG, "dlu.py"))
workloads = _load("dlu_workloads", os.path.join(PKG, "workloads.py"))


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.asfortranarray(a, dtype=np.float64).tobytes(order="F")).hexdigest()


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def reconstruction_error(original, factor, ipiv, jpiv, info):
    """max |P A Q - Lbar Ubar| (delayed_lu_cuda/example.cu:24-80), in numpy."""
    m, n = original.shape
    e = info
    lo = np.eye(m)
    up = np.zeros((m, n))
    for k in range(e):
        lo[k + 1:, k] = factor[k + 1:, k]
        up[k, k:] = factor[k, k:]
    up[e:, e:] = factor[e:, e:]
    rows = oracle.replay(ipiv, e, m)
    cols = oracle.replay(jpiv, e, n)
    return np.abs(lo @ up - original[np.ix_(rows, cols)]).max()


def delayed_set(jpiv, info, n):
    """Original column indices that end up in the delayed positions [info, n)."""
    order = oracle.replay(jpiv, info, n)
    return np.sort(order[info:])
