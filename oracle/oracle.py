"""CPU oracle for the delayed-pivoting LU -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import this
module, and only as the checker (or the timed CPU baseline).  The product library never loads it.

It wraps ``oracle/_build/libdlu_oracle.so`` (built from ``dlu_oracle.c`` with gcc), which restates
the reference's ``delayed_getrf`` (delayed_lu/delayed_dgetrf.cpp:129-251) and ``lapack_getrf``
(delayed_lu/lapack_dgetrf.cpp:21-63) on top of a BLAS/LAPACK found at run time.
"""
from __future__ import annotations

import ctypes
import glob
import os
import subprocess
import sys
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libdlu_oracle.so")
SOURCES = [os.path.join(HERE, "dlu_oracle.c")]

_lib = None


def build(force: bool = False, out: str = LIB_PATH, defines=()) -> str:
    """Compile the oracle with gcc (no BLAS at link time; it is opened at run time).

    ``defines`` (e.g. ``["DLU_ORACLE_MUTANT_TIES"]``) and ``out`` serve the tests' planted-mistake
    build, which goes to a path of its own."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    newest = max(os.path.getmtime(s) for s in SOURCES + [os.path.join(HERE, "dlu_oracle.h")])
    if not force and os.path.exists(out) and os.path.getmtime(out) >= newest:
        return out
    tmp = out + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
           "-o", tmp] + [f"-D{d}" for d in defines] + SOURCES + ["-ldl", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def load_build(path: str):
    """Open another build of the oracle (same BLAS as lib()) for delayed_getrf(library=...)."""
    lib()
    L = ctypes.CDLL(path)
    L.dlu_oracle_init.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    for name in ("dlu_oracle_delayed_getrf", "dlu_oracle_delayed_getrf_column"):
        getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, ip, ip,
                                     ctypes.c_int, dp, ip]
    sy = [ctypes.c_int, dp, ctypes.c_int, ip, ip, ip]
    L.dlu_oracle_delayed_sytrf.argtypes = sy
    L.dlu_oracle_delayed_sytrf_blocked.argtypes = sy + [ctypes.c_int]
    L.dlu_oracle_sytrf_unblocked_ld.argtypes = sy
    for path_, prefix in _blas_candidates():
        if L.dlu_oracle_init(path_.encode(), prefix.encode()) == 0:
            return L
    raise RuntimeError("oracle build: no usable BLAS/LAPACK found")


def _blas_candidates():
    env = os.environ.get("DLU_ORACLE_BLAS")
    if env:
        path, _, prefix = env.partition(":")
        yield path, prefix
    for p in ("/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"):
        if os.path.exists(p):
            yield p, ""
    try:
        import scipy  # noqa: F401  (only to locate its bundled OpenBLAS)
        libs = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        for p in sorted(glob.glob(os.path.join(libs, "libscipy_openblas-*.so"))):
            yield p, "scipy_"
    except ImportError:
        pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    # MKL: share the GNU OpenMP runtime that torch already uses instead of loading libiomp5.
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    L = ctypes.CDLL(LIB_PATH)
    L.dlu_oracle_init.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    L.dlu_oracle_init.restype = ctypes.c_int
    L.dlu_oracle_blas_name.restype = ctypes.c_char_p
    L.dlu_oracle_set_threads.argtypes = [ctypes.c_int]
    L.dlu_oracle_set_threads.restype = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    L.dlu_oracle_delayed_getrf.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, ip,
                                           ip, ctypes.c_int, dp, ip]
    L.dlu_oracle_delayed_getrf_column.argtypes = L.dlu_oracle_delayed_getrf.argtypes
    L.dlu_oracle_lapack_getrf.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, ip]
    L.dlu_oracle_unblocked_column_ld.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int,
                                                 ip, ip, ip]
    L.dlu_oracle_unblocked_ld.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, ip,
                                          ip]
    L.dlu_oracle_backward_error.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int,
                                            ip, ip, ctypes.c_int, ctypes.c_int]
    L.dlu_oracle_backward_error.restype = ctypes.c_double
    sy = [ctypes.c_int, dp, ctypes.c_int, ip, ip, ip]
    L.dlu_oracle_delayed_sytrf.argtypes = sy
    L.dlu_oracle_delayed_sytrf_blocked.argtypes = sy + [ctypes.c_int]
    L.dlu_oracle_sytrf_unblocked_ld.argtypes = sy
    errors = []
    for path, prefix in _blas_candidates():
        rc = L.dlu_oracle_init(path.encode(), prefix.encode())
        if rc == 0:
            break
        errors.append(f"{path}: rc={rc}")
    else:
        raise RuntimeError("oracle: no usable BLAS/LAPACK found: " + "; ".join(errors))
    _lib = L
    return L


def blas_name() -> str:
    return lib().dlu_oracle_blas_name().decode()


def set_threads(n: int) -> int:
    return lib().dlu_oracle_set_threads(int(n))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if a is not None else None


def _fortran(a: np.ndarray) -> np.ndarray:
    return np.array(a, dtype=np.float64, order="F", copy=True)


def delayed_getrf(a: np.ndarray, nb: int = 64, margins: bool = False, column: bool = False,
                  library=None):
    """Restated delayed_getrf (``column=True``: delayed_getrf_column) on a copy of ``a``.

    Returns ``(lu, ipiv, jpiv, info, extra)`` where ``extra`` holds ``fallbacks`` and, when
    requested, the per-step column-pivot ``margins``.  ``library`` may name another build of the
    oracle (the tests' planted-mistake build); by default the one ``build()`` makes.
    """
    L = library or lib()
    fn = L.dlu_oracle_delayed_getrf_column if column else L.dlu_oracle_delayed_getrf
    f = _fortran(a)
    m, n = f.shape
    p = min(m, n)
    ipiv = np.zeros(max(p, 1), dtype=np.int32)
    jpiv = np.zeros(max(p, 1), dtype=np.int32)
    info = ctypes.c_int(-1)
    fb = ctypes.c_int(0)
    marg = np.zeros(max(p, 1), dtype=np.float64) if margins else None
    fn(m, n, _dp(f), max(1, m), _ip(ipiv), _ip(jpiv), ctypes.byref(info), nb, _dp(marg),
       ctypes.byref(fb))
    extra = {"fallbacks": fb.value}
    if margins:
        extra["margins"] = marg[:p]
    return f, ipiv[:p], jpiv[:p], info.value, extra


def lapack_getrf(a: np.ndarray):
    f = _fortran(a)
    m, n = f.shape
    p = min(m, n)
    ipiv = np.zeros(max(p, 1), dtype=np.int32)
    info = ctypes.c_int(-1)
    lib().dlu_oracle_lapack_getrf(m, n, _dp(f), max(1, m), _ip(ipiv), ctypes.byref(info))
    return f, ipiv[:p], info.value


def lapack_sgetrf(a: np.ndarray):
    """LAPACK sgetrf on a float32 copy: (lu32, ipiv 0-based, info)."""
    L = lib()
    f = np.array(a, dtype=np.float32, order="F", copy=True)
    m, n = f.shape
    p = min(m, n)
    ipiv = np.zeros(max(p, 1), dtype=np.int32)
    info = ctypes.c_int(-1)
    L.dlu_oracle_lapack_sgetrf.argtypes = [ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int),
                                           ctypes.POINTER(ctypes.c_int)]
    rc = L.dlu_oracle_lapack_sgetrf(m, n, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                    max(1, m), _ip(ipiv), ctypes.byref(info))
    if rc != 0:
        raise RuntimeError("the loaded LAPACK has no sgetrf_")
    return f, ipiv[:p], info.value


def unblocked_ld(a: np.ndarray, column: bool = False):
    """The independent unblocked long-double restatement (no BLAS), either variant."""
    f = _fortran(a)
    m, n = f.shape
    p = min(m, n)
    ipiv = np.zeros(max(p, 1), dtype=np.int32)
    jpiv = np.zeros(max(p, 1), dtype=np.int32)
    info = ctypes.c_int(-1)
    fn = lib().dlu_oracle_unblocked_column_ld if column else lib().dlu_oracle_unblocked_ld
    fn(m, n, _dp(f), max(1, m), _ip(ipiv), _ip(jpiv), ctypes.byref(info))
    return f, ipiv[:p], jpiv[:p], info.value


def backward_error(original: np.ndarray, factor: np.ndarray, ipiv, jpiv: Optional[np.ndarray],
                   info: int, delayed: bool = True) -> float:
    """||P A Q - Lbar Ubar||_F / (max(m,n) eps ||A||_F) of a compact factorization."""
    o = _fortran(original)
    f = _fortran(factor)
    m, n = o.shape
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    jp = np.ascontiguousarray(jpiv, dtype=np.int32) if jpiv is not None else None
    return lib().dlu_oracle_backward_error(m, n, _dp(o), _dp(f), max(1, m), _ip(ip), _ip(jp),
                                           int(info), 1 if delayed else 0)


def delayed_sytrf(a: np.ndarray, blocked: bool = True, nb: int = 64, ld: bool = False,
                  library=None):
    """Delayed Bunch-Kaufman LDL^T of the symmetric matrix in the lower triangle of `a` (a copy):
    blocked (delayed_sytrf_blocked.cpp, panels of nb), unblocked (delayed_sytrf.cpp), or the
    independent long-double restatement (ld=True).  Returns (factor, perm, piv_size, info)."""
    f = _fortran(a)
    n = f.shape[0]
    perm = np.zeros(max(n, 1), dtype=np.int32)
    piv = np.zeros(max(n, 1), dtype=np.int32)
    info = ctypes.c_int(-1)
    L = library or lib()
    if ld:
        L.dlu_oracle_sytrf_unblocked_ld(n, _dp(f), max(1, n), _ip(perm), _ip(piv),
                                        ctypes.byref(info))
    elif blocked:
        L.dlu_oracle_delayed_sytrf_blocked(n, _dp(f), max(1, n), _ip(perm), _ip(piv),
                                           ctypes.byref(info), nb)
    else:
        L.dlu_oracle_delayed_sytrf(n, _dp(f), max(1, n), _ip(perm), _ip(piv), ctypes.byref(info))
    if info.value < -1:
        raise RuntimeError("oracle: the BLAS lacks a routine of the symmetric factorization")
    return f, perm[:n], piv[:n], info.value


def sytrf_parts(factor: np.ndarray, piv_size, info: int):
    """Unit lower L and block-diagonal D of a delayed LDL^T factor (delayed-bunch-kaufman/
    example.cpp:47-148): L's column of a 1x1 pivot holds its multipliers below the diagonal, a
    2x2 pivot's two columns hold theirs below the block; D has the 1x1 values, the symmetric 2x2
    blocks, and the delayed part's Schur complement as its last block (rows/columns [info, n),
    read symmetrically from the lower triangle)."""
    f = np.asarray(factor, dtype=np.float64)
    n = f.shape[0]
    L = np.eye(n)
    D = np.zeros((n, n))
    k = 0
    while k < info:
        if piv_size[k] == 2:
            D[k, k], D[k + 1, k + 1] = f[k, k], f[k + 1, k + 1]
            D[k + 1, k] = D[k, k + 1] = f[k + 1, k]
            L[k + 2:, k] = f[k + 2:, k]
            L[k + 2:, k + 1] = f[k + 2:, k + 1]
            k += 2
        else:
            D[k, k] = f[k, k]
            L[k + 1:, k] = f[k + 1:, k]
            k += 1
    s = np.tril(f[info:, info:])
    D[info:, info:] = s + np.tril(s, -1).T
    return L, D


def sytrf_reconstruction_error(original: np.ndarray, factor: np.ndarray, perm, piv_size,
                               info: int) -> float:
    """max |P^T A P - L D L^T| relative to max |A| (A read symmetrically from its lower
    triangle), as delayed-bunch-kaufman/example.cpp:97-148 checks it."""
    o = np.tril(np.asarray(original, dtype=np.float64))
    a = o + np.tril(o, -1).T
    L, D = sytrf_parts(factor, piv_size, info)
    p = np.asarray(perm)
    err = np.abs(a[np.ix_(p, p)] - L @ D @ L.T).max()
    scale = np.abs(a).max()
    return float(err / scale) if scale > 0 else float(err)


def replay(piv, count: int, size: int) -> np.ndarray:
    """Order vector from a swap history: swap(order[k], order[piv[k]]) for k < count."""
    order = np.arange(size)
    for k in range(count):
        t = order[k]
        order[k] = order[piv[k]]
        order[piv[k]] = t
    return order


if __name__ == "__main__":  # manual probe: python oracle/oracle.py
    build(force="--force" in sys.argv)
    print(blas_name())


def front_factor(front: np.ndarray, ns: int, nb: int = 64):
    """Steps 1-2 of a front (doc/（非）对称稠密矩阵的延迟主元LU方法.md:228-317), restated on a copy:
    delayed_getrf of the leading ns x ns fully summed block (Step 1), its interchanges extended
    across the front, then U_EU = L_EE^-1 A_EU, L_UE = A_UE U_EE^-1, A_DU -= L_DE U_EU,
    A_UD -= L_UE U_ED and A_UU -= L_UE U_EU (Step 2).  Returns (F, ipiv, jpiv, info)."""
    import scipy.linalg
    f = _fortran(front)
    nf = f.shape[0]
    lu, ipiv, jpiv, e, _ = delayed_getrf(f[:ns, :ns], nb=nb)
    f[:ns, :ns] = lu
    if nf > ns and e > 0:
        f[:ns, ns:] = f[:ns, ns:][replay(ipiv, e, ns), :]
        f[ns:, :ns] = f[ns:, :ns][:, replay(jpiv, e, ns)]
        l_ee = np.tril(f[:e, :e], -1) + np.eye(e)
        u_ee = np.triu(f[:e, :e])
        f[:e, ns:] = scipy.linalg.solve_triangular(l_ee, f[:e, ns:], lower=True,
                                                   unit_diagonal=True)
        f[ns:, :e] = scipy.linalg.solve_triangular(u_ee.T, f[ns:, :e].T, lower=True).T
        f[e:, ns:] -= f[e:, :e] @ f[:e, ns:]
        f[ns:, e:ns] -= f[ns:, :e] @ f[:e, e:ns]
    return f, ipiv, jpiv, e
