"""ctypes binding of libdlu.so (the C ABI declared in include/dlu.h).

Mirrors the reference's GPU interface (delayed_lu_cuda/delayed_dgetrf_cuda.hpp:7-63): the same
argument meaning, status ordinals and ``info`` rules.  Device buffers come from PyTorch (plumbing
only); the factorization itself always runs in libdlu's HIP kernels.  There is no CPU fallback:
if the library is missing or a GPU call fails, these functions raise.
"""
from __future__ import annotations

import atexit
import ctypes
import enum
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libdlu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "dlu.h")


class Status(enum.IntEnum):
    SUCCESS = 0
    INVALID_ARGUMENT = 1
    ALLOCATION_FAILED = 2
    RUNTIME_FAILURE = 3
    KERNEL_FAILURE = 4


class Options(ctypes.Structure):
    _fields_ = [("nb", ctypes.c_int), ("lookahead", ctypes.c_int),
                ("spin_limit_ticks", ctypes.c_longlong), ("profile", ctypes.c_int),
                ("panel_cus", ctypes.c_int), ("rows_per_lane", ctypes.c_int),
                ("merge_rows", ctypes.c_int), ("split_update", ctypes.c_int),
                ("fallback", ctypes.c_int), ("panel_kernel", ctypes.c_int),
                ("defer_left", ctypes.c_int)]


PANEL_AUTO = 0  # one exchange per pivot step where the panel fits it
PANEL_MOVE = 1  # the exchange-and-move panel kernel (two exchanges per step)
PANEL_LAG = 2   # one exchange per step, or an error


FALLBACK_FULL = 0    # delayed_getrf: largest |a| of the active block
FALLBACK_COLUMN = 1  # delayed_getrf_column: rightmost nonzero column


class Report(ctypes.Structure):
    _fields_ = [("panels", ctypes.c_int), ("fallbacks", ctypes.c_int),
                ("near_ties", ctypes.c_int), ("exact_ties", ctypes.c_int),
                ("min_margin", ctypes.c_double),
                ("update_launches", ctypes.c_int), ("update_seconds", ctypes.c_double),
                ("update_flops", ctypes.c_double), ("large_update_launches", ctypes.c_int),
                ("large_update_seconds", ctypes.c_double), ("large_update_flops", ctypes.c_double),
                ("panel_launches", ctypes.c_int),
                ("panel_seconds", ctypes.c_double), ("solve_launches", ctypes.c_int),
                ("solve_seconds", ctypes.c_double),
                ("panel_phase_seconds", ctypes.c_double * 8),
                ("large_update_bytes", ctypes.c_double), ("search_launches", ctypes.c_int),
                ("search_seconds", ctypes.c_double), ("search_bytes", ctypes.c_double)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["panel_phase_seconds"] = list(self.panel_phase_seconds)
        return d


class MixedOptions(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iter", ctypes.c_int), ("nb", ctypes.c_int),
                ("fp64_fallback", ctypes.c_int)]


class MixedReport(ctypes.Structure):
    _fields_ = [("factor_info", ctypes.c_int), ("iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("used_fp64", ctypes.c_int),
                ("factor_seconds", ctypes.c_double), ("refine_seconds", ctypes.c_double),
                ("total_seconds", ctypes.c_double), ("max_backward_error", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BatchItem(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("lda", ctypes.c_int), ("lu", ctypes.c_void_p),
                ("ldlu", ctypes.c_int), ("ipiv", ctypes.c_void_p), ("jpiv", ctypes.c_void_p),
                ("info", ctypes.c_void_p), ("info_in", ctypes.c_int), ("n", ctypes.c_int),
                ("nrhs", ctypes.c_int), ("b", ctypes.c_void_p), ("ldb", ctypes.c_int),
                ("x", ctypes.c_void_p), ("ldx", ctypes.c_int), ("solve_only", ctypes.c_int)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_gib", ctypes.c_double), ("batch_cutoff", ctypes.c_int),
                ("mixed", ctypes.c_int), ("tol", ctypes.c_double), ("max_batch", ctypes.c_int),
                ("large_workers", ctypes.c_int), ("eager_sweeps", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("n", ctypes.c_int), ("nrhs", ctypes.c_int),
                ("a", ctypes.c_void_p), ("lda", ctypes.c_int), ("b", ctypes.c_void_p),
                ("ldb", ctypes.c_int), ("x", ctypes.c_void_p), ("ldx", ctypes.c_int),
                ("factor_of", ctypes.c_longlong), ("keep", ctypes.c_int)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("info", ctypes.c_int),
                ("path", ctypes.c_int), ("iterations", ctypes.c_int), ("reused", ctypes.c_int),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double),
                ("t_done", ctypes.c_double), ("backward_error", ctypes.c_double),
                ("graph_sweeps", ctypes.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    """Load libdlu.so; raise if it has not been built (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libdlu.so is not built ({LIB_PATH}); run __graft_entry__.build()")
    # One HIP runtime per process: PyTorch ships its own libamdhip64, and a process that loads
    # libdlu first (which would bring in the system's) cannot then use the GPU through both.  When
    # PyTorch is installed, it is loaded first and libdlu binds to its runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    sig = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, _ip, _ip, _ip,
           ctypes.c_void_p]
    for name in ("dlu_getrf_device", "dlu_getrf_host", "dlu_getrf_native_device",
                 "dlu_getrf_native_host", "dlu_getrf_hybrid_device", "dlu_getrf_hybrid_host",
                 "dlu_getrf_column_device", "dlu_getrf_column_host", "dlu_getrf_device_noalloc"):
        fn = getattr(L, name)
        fn.argtypes = sig
        fn.restype = ctypes.c_int
    L.dlu_getrf_device_ex.argtypes = sig + [ctypes.POINTER(Options), ctypes.POINTER(Report), _dp]
    L.dlu_getrf_device_ex.restype = ctypes.c_int
    L.dlu_status_string.argtypes = [ctypes.c_int]
    L.dlu_status_string.restype = ctypes.c_char_p
    L.dlu_gemm_sub_device.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 3 + [
        ctypes.c_int, ctypes.c_void_p]
    L.dlu_gemm_sub_device.restype = ctypes.c_int
    L.dlu_gemm_sub_device_config.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 3 + [
        ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.dlu_gemm_sub_device_config.restype = ctypes.c_int
    L.dlu_tuned_gemm_config.argtypes = [ctypes.c_int] * 3
    L.dlu_tuned_gemm_config.restype = ctypes.c_int
    L.dlu_tuned_batch_cutoff.argtypes = []
    L.dlu_tuned_batch_cutoff.restype = ctypes.c_int
    L.dlu_sgetrf_device.argtypes = sig
    L.dlu_sgetrf_device.restype = ctypes.c_int
    L.dlu_sgetrf_device_ex.argtypes = sig + [ctypes.POINTER(Options), ctypes.POINTER(Report), _dp]
    L.dlu_sgetrf_device_ex.restype = ctypes.c_int
    rs = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, _ip, _ip, ctypes.c_int,
          ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    for name in ("dlu_getrs_device", "dlu_sgetrs_device"):
        getattr(L, name).argtypes = rs
        getattr(L, name).restype = ctypes.c_int
    L.dlu_backward_error_device.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_void_p]
    L.dlu_backward_error_device.restype = ctypes.c_int
    gs = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
          ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
          ctypes.POINTER(MixedOptions), ctypes.POINTER(MixedReport), _dp]
    for name in ("dlu_gesv_device", "dlu_gesv_mixed_device"):
        getattr(L, name).argtypes = gs
        getattr(L, name).restype = ctypes.c_int
    L.dlu_batch_gesv_device.argtypes = [ctypes.c_int, ctypes.POINTER(BatchItem), ctypes.c_void_p]
    L.dlu_batch_gesv_device.restype = ctypes.c_int
    L.dlu_server_create.argtypes = [ctypes.POINTER(ServerConfig)]
    L.dlu_server_create.restype = ctypes.c_void_p
    L.dlu_server_batch_cutoff.argtypes = [ctypes.c_void_p]
    L.dlu_server_batch_cutoff.restype = ctypes.c_int
    L.dlu_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request)]
    L.dlu_server_submit.restype = ctypes.c_int
    L.dlu_server_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(Completion), ctypes.c_int]
    L.dlu_server_poll.restype = ctypes.c_int
    L.dlu_server_drain.argtypes = [ctypes.c_void_p]
    L.dlu_server_drain.restype = ctypes.c_int
    L.dlu_server_now.argtypes = [ctypes.c_void_p]
    L.dlu_server_now.restype = ctypes.c_double
    L.dlu_server_destroy.argtypes = [ctypes.c_void_p]
    L.dlu_server_destroy.restype = None
    L.dlu_front_factor_device.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_int, _ip, _ip, _ip, ctypes.c_void_p]
    L.dlu_front_factor_device.restype = ctypes.c_int
    L.dlu_extend_add_device.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.dlu_extend_add_device.restype = ctypes.c_int
    L.dlu_shutdown.argtypes = []
    L.dlu_shutdown.restype = None
    sy = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, _ip, _ip, _ip]
    L.dlu_sytrf_device.argtypes = sy + [ctypes.c_void_p]
    L.dlu_sytrf_device.restype = ctypes.c_int
    L.dlu_sytrf_host.argtypes = sy + [ctypes.c_void_p]
    L.dlu_sytrf_host.restype = ctypes.c_int
    L.dlu_sytrf_device_nb.argtypes = sy + [ctypes.c_int, ctypes.c_void_p]
    L.dlu_sytrf_device_nb.restype = ctypes.c_int
    _lib = L
    # Free libdlu's cached device state while the HIP runtime is still up (atexit handlers run
    # before the interpreter and the runtime's own teardown).
    atexit.register(shutdown)
    return L


def shutdown() -> None:
    """dlu_shutdown(): release libdlu's cached workspaces and this thread's solve scratch."""
    if _lib is not None:
        _lib.dlu_shutdown()


def net_order(piv, steps: int, n: int) -> np.ndarray:
    """Original index at each position after the swap history piv[0:steps] (0-based swaps)."""
    order = np.arange(n)
    for k in range(steps):
        p = int(piv[k])
        order[k], order[p] = order[p], order[k]
    return order


def batch_gesv_device(items, stream=None) -> Status:
    """One batched launch over a list of BatchItem (host structures holding device pointers)."""
    arr = (BatchItem * max(len(items), 1))(*items)
    return Status(lib().dlu_batch_gesv_device(len(items), arr, _stream_ptr(stream)))


class Server:
    """Thin wrapper of dlu_server_*: submit requests, poll completions, drain, close."""

    def __init__(self, budget_gib=32.0, batch_cutoff=0, mixed=False, tol=1e-13, max_batch=1024,
                 large_workers=0, graphs=True):
        cfg = ServerConfig(budget_gib, batch_cutoff, 1 if mixed else 0, tol, max_batch,
                           large_workers, 0 if graphs else 1)
        self._h = lib().dlu_server_create(ctypes.byref(cfg))
        if not self._h:
            raise RuntimeError("dlu_server_create failed")

    @property
    def cutoff(self) -> int:
        return lib().dlu_server_batch_cutoff(self._h)

    def now(self) -> float:
        return lib().dlu_server_now(self._h)

    def submit(self, rid, n, nrhs, a, lda, b, ldb, x, ldx, factor_of=-1, keep=False) -> Status:
        req = Request(rid, n, nrhs, a, lda, b, ldb, x, ldx, factor_of, 1 if keep else 0)
        return Status(lib().dlu_server_submit(self._h, ctypes.byref(req)))

    def poll(self, max_items=4096):
        buf = (Completion * max_items)()
        k = lib().dlu_server_poll(self._h, buf, max_items)
        return [buf[i].as_dict() for i in range(k)]

    def drain(self) -> Status:
        return Status(lib().dlu_server_drain(self._h))

    def close(self):
        if self._h:
            lib().dlu_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def status_string(status: int) -> str:
    return lib().dlu_status_string(int(status)).decode()


def declared_symbols() -> list:
    """Function names declared in include/dlu.h (for the export check)."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dlu_\w+)\s*\(", text, re.M)))


def _stream_ptr(stream) -> Optional[int]:
    if stream is None:
        return None
    return int(getattr(stream, "cuda_stream", stream))


def getrf_device(m: int, n: int, d_a: int, lda: int, stream=None, options: Optional[dict] = None,
                 margins: bool = False, variant: str = "device"):
    """Factor the device matrix at address ``d_a`` in place.

    Returns ``(status, ipiv, jpiv, info, report, margins)``; ``report`` and ``margins`` are None
    unless the extended entry point is used (``options`` given or ``margins=True``).
    """
    L = lib()
    p = max(min(m, n), 0)
    ipiv = np.zeros(max(p, 1), dtype=np.int32)
    jpiv = np.zeros(max(p, 1), dtype=np.int32)
    info = ctypes.c_int(-7)
    ip = ipiv.ctypes.data_as(_ip)
    jp = jpiv.ctypes.data_as(_ip)
    s = _stream_ptr(stream)
    if options is not None or margins:
        opt = Options(**(options or {}))
        rep = Report()
        marg = np.zeros(max(p, 1), dtype=np.float64) if margins else None
        st = L.dlu_getrf_device_ex(m, n, d_a, lda, ip, jp, ctypes.byref(info), s,
                                   ctypes.byref(opt), ctypes.byref(rep),
                                   marg.ctypes.data_as(_dp) if margins else None)
        return (Status(st), ipiv[:p], jpiv[:p], info.value, rep.as_dict(),
                marg[:p] if margins else None)
    fn = {"device": L.dlu_getrf_device, "native": L.dlu_getrf_native_device,
          "hybrid": L.dlu_getrf_hybrid_device, "column": L.dlu_getrf_column_device}[variant]
    st = fn(m, n, d_a, lda, ip, jp, ctypes.byref(info), s)
    return Status(st), ipiv[:p], jpiv[:p], info.value, None, None


def sgetrf_device(m: int, n: int, d_a: int, lda: int, stream=None,
                  options: Optional[dict] = None, margins: bool = False):
    """FP32 factorization of the float32 device matrix at ``d_a``; same returns as getrf_device."""
    L = lib()
    p = max(min(m, n), 0)
    ipiv = np.zeros(max(p, 1), dtype=np.int32)
    jpiv = np.zeros(max(p, 1), dtype=np.int32)
    info = ctypes.c_int(-7)
    s = _stream_ptr(stream)
    opt = Options(**(options or {}))
    rep = Report()
    marg = np.zeros(max(p, 1), dtype=np.float64) if margins else None
    st = L.dlu_sgetrf_device_ex(m, n, d_a, lda, ipiv.ctypes.data_as(_ip),
                                jpiv.ctypes.data_as(_ip), ctypes.byref(info), s,
                                ctypes.byref(opt), ctypes.byref(rep),
                                marg.ctypes.data_as(_dp) if margins else None)
    return (Status(st), ipiv[:p], jpiv[:p], info.value, rep.as_dict(),
            marg[:p] if margins else None)


def sytrf_device(n: int, d_a: int, lda: int, stream=None, nb: int = 64):
    """Delayed Bunch-Kaufman LDL^T of the symmetric device matrix (lower triangle) at ``d_a``:
    (status, perm, piv_size, info)."""
    L = lib()
    perm = np.zeros(max(n, 1), dtype=np.int32)
    piv = np.zeros(max(n, 1), dtype=np.int32)
    info = ctypes.c_int(-7)
    st = L.dlu_sytrf_device_nb(n, d_a, lda, perm.ctypes.data_as(_ip), piv.ctypes.data_as(_ip),
                               ctypes.byref(info), nb, _stream_ptr(stream))
    return Status(st), perm[:n], piv[:n], info.value


def getrs_device(n: int, nrhs: int, d_lu: int, lda: int, ipiv, jpiv, info: int, d_b: int,
                 ldb: int, stream=None, single: bool = False) -> Status:
    """Solve with factors from getrf_device (or sgetrf_device with ``single=True``); d_b := X."""
    ip = np.ascontiguousarray(ipiv, dtype=np.int32)
    jp = np.ascontiguousarray(jpiv, dtype=np.int32)
    fn = lib().dlu_sgetrs_device if single else lib().dlu_getrs_device
    return Status(fn(n, nrhs, d_lu, lda, ip.ctypes.data_as(_ip), jp.ctypes.data_as(_ip), info,
                     d_b, ldb, _stream_ptr(stream)))


def backward_error_device(n: int, nrhs: int, d_a: int, lda: int, d_x: int, ldx: int, d_b: int,
                          ldb: int, stream=None):
    """Per-column ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf): (status, array)."""
    out = np.zeros(max(nrhs, 1), dtype=np.float64)
    st = lib().dlu_backward_error_device(n, nrhs, d_a, lda, d_x, ldx, d_b, ldb,
                                         out.ctypes.data_as(_dp), _stream_ptr(stream))
    return Status(st), out[:nrhs]


def gesv_device(n: int, nrhs: int, d_a: int, lda: int, d_b: int, ldb: int, d_x: int, ldx: int,
                stream=None, options: Optional[dict] = None, mixed: bool = True):
    """A X = B (FP64 in and out): mixed precision (FP32 factor + FP64 refinement) or plain FP64.

    Returns (status, report dict, per-column backward errors)."""
    opt = MixedOptions(**(options or {}))
    rep = MixedReport()
    bwd = np.zeros(max(nrhs, 1), dtype=np.float64)
    fn = lib().dlu_gesv_mixed_device if mixed else lib().dlu_gesv_device
    st = fn(n, nrhs, d_a, lda, d_b, ldb, d_x, ldx, _stream_ptr(stream), ctypes.byref(opt),
            ctypes.byref(rep), bwd.ctypes.data_as(_dp))
    return Status(st), rep.as_dict(), bwd[:nrhs]


def getrf_host(a: np.ndarray, variant: str = "host", stream=None):
    """Host-memory interface on a Fortran-ordered copy of ``a``: (status, lu, ipiv, jpiv, info)."""
    L = lib()
    f = np.array(a, dtype=np.float64, order="F", copy=True)
    m, n = f.shape
    p = min(m, n)
    ipiv = np.zeros(max(p, 1), dtype=np.int32)
    jpiv = np.zeros(max(p, 1), dtype=np.int32)
    info = ctypes.c_int(-7)
    fn = {"host": L.dlu_getrf_host, "native": L.dlu_getrf_native_host,
          "hybrid": L.dlu_getrf_hybrid_host, "column": L.dlu_getrf_column_host}[variant]
    st = fn(m, n, f.ctypes.data, max(1, m), ipiv.ctypes.data_as(_ip), jpiv.ctypes.data_as(_ip),
            ctypes.byref(info), _stream_ptr(stream))
    return Status(st), f, ipiv[:p], jpiv[:p], info.value


def front_factor_device(nf: int, ns: int, d_f: int, ldf: int, stream=None):
    """Steps 1-2 of a front (dlu_front_factor_device): (status, ipiv, jpiv, info), the pivots
    being the fully summed block's swap histories (ns entries)."""
    ipiv = np.zeros(max(ns, 1), dtype=np.int32)
    jpiv = np.zeros(max(ns, 1), dtype=np.int32)
    info = ctypes.c_int(-7)
    st = lib().dlu_front_factor_device(nf, ns, d_f, ldf, ipiv.ctypes.data_as(_ip),
                                       jpiv.ctypes.data_as(_ip), ctypes.byref(info),
                                       _stream_ptr(stream))
    return Status(st), ipiv[:ns], jpiv[:ns], info.value


def extend_add_device(rows: int, cols: int, d_child: int, ldc: int, d_map_r: int, d_map_c: int,
                      d_parent: int, ldp: int, stream=None) -> Status:
    """parent(map_r[i], map_c[j]) += child(i, j) (dlu_extend_add_device; asynchronous)."""
    return Status(lib().dlu_extend_add_device(rows, cols, d_child, ldc, d_map_r, d_map_c, d_parent,
                                              ldp, _stream_ptr(stream)))


def contribution_indices(row_ids, col_ids, ns: int, ipiv, jpiv, info: int):
    """Global row and column indices of a front's contribution block F[info:nf, info:nf] after
    dlu_front_factor_device (doc Step 3.1: the delayed variables D, then the update variables U,
    rows and columns kept apart for the unsymmetric case).  row_ids / col_ids are the front's
    global indices before factoring (fully summed ones first)."""
    rows = np.array(row_ids, dtype=np.int64)
    cols = np.array(col_ids, dtype=np.int64)
    rows[:ns] = rows[:ns][net_order(ipiv, info, ns)]
    cols[:ns] = cols[:ns][net_order(jpiv, info, ns)]
    return rows[info:], cols[info:]


def parent_positions(parent_ids, child_ids):
    """Positions of a child's global indices in a parent front's index list (doc Step 3.2-3.3):
    indices the parent lacks (the child's delayed variables) are appended to it.  Returns
    (new parent index list, positions)."""
    ids = list(int(g) for g in parent_ids)
    where = {g: i for i, g in enumerate(ids)}
    pos = []
    for g in child_ids:
        g = int(g)
        if g not in where:
            where[g] = len(ids)
            ids.append(g)
        pos.append(where[g])
    return np.array(ids, dtype=np.int64), np.array(pos, dtype=np.int32)


def gemm_sub_device(M, N, K, c_ptr, a_ptr, b_ptr, ld, stream=None, config=-1) -> Status:
    """C -= A B on the trailing-update kernel (config: tile configuration, -1 = tuned)."""
    return Status(lib().dlu_gemm_sub_device_config(M, N, K, c_ptr, a_ptr, b_ptr, ld, config,
                                                   _stream_ptr(stream)))


# ---- torch plumbing: column-major device matrices -------------------------------------------

def to_device(a: np.ndarray, lda: Optional[int] = None):
    """Copy a host matrix to a column-major CUDA/HIP tensor; returns (storage tensor, lda).

    The storage tensor has shape (n, lda): row j is column j of the matrix.
    """
    import torch
    m, n = a.shape
    lda = lda or max(1, m)
    host = np.zeros((n, lda), dtype=np.float64)
    host[:, :m] = np.asarray(a, dtype=np.float64).T
    return torch.from_numpy(host).to("cuda"), lda


def from_device(t, m: int) -> np.ndarray:
    """Inverse of to_device: the m x n matrix held by storage tensor t (shape (n, lda))."""
    return np.asfortranarray(t.cpu().numpy()[:, :m].T)
