"""CPU checks of the C ABI: the library loads, exports every declared entry point, and rejects
invalid arguments before touching the GPU (delayed_lu_cuda/delayed_dgetrf_cuda.cu:385-397,
403-409, 449-460)."""
import ctypes
import subprocess

import numpy as np
import pytest

from helpers import dlu


def test_exports_every_declared_symbol():
    declared = dlu.declared_symbols()
    assert "dlu_getrf_device" in declared and "dlu_status_string" in declared
    out = subprocess.run(["nm", "-D", "--defined-only", dlu.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    L = dlu.lib()
    for s in declared:
        assert hasattr(L, s)


def test_status_strings_and_ordinals():
    assert [int(s) for s in dlu.Status] == [0, 1, 2, 3, 4]
    assert dlu.status_string(0) == "success"
    assert dlu.status_string(1) == "invalid argument"
    assert dlu.status_string(99) == "unknown status"


@pytest.mark.parametrize("m,n,lda", [(-1, 3, 3), (3, -1, 3), (3, 3, 2), (0, 0, 0)])
def test_invalid_dimensions(m, n, lda):
    ip = np.zeros(4, dtype=np.int32)
    info = ctypes.c_int(7)
    st = dlu.lib().dlu_getrf_device(m, n, 1234, lda, ip.ctypes.data_as(dlu._ip),
                                    ip.ctypes.data_as(dlu._ip), ctypes.byref(info), None)
    assert st == dlu.Status.INVALID_ARGUMENT and info.value == -1


def test_null_pointers():
    ip = np.zeros(4, dtype=np.int32)
    info = ctypes.c_int(7)
    L = dlu.lib()
    # null matrix with m, n > 0
    assert L.dlu_getrf_device(3, 3, None, 3, ip.ctypes.data_as(dlu._ip),
                              ip.ctypes.data_as(dlu._ip), ctypes.byref(info), None) == 1
    assert info.value == -1
    # null ipiv / jpiv with min(m, n) > 0
    info.value = 7
    assert L.dlu_getrf_device(3, 3, 1234, 3, None, ip.ctypes.data_as(dlu._ip),
                              ctypes.byref(info), None) == 1
    assert info.value == -1
    info.value = 7
    assert L.dlu_getrf_host(3, 3, None, 3, ip.ctypes.data_as(dlu._ip),
                            ip.ctypes.data_as(dlu._ip), ctypes.byref(info), None) == 1
    assert info.value == -1


def test_invalid_options_rejected_before_device_work():
    # an unsupported panel width, rows per lane or fallback variant is rejected (info = -1)
    # before any device call, so this runs without a GPU
    for opts in ({"nb": 96}, {"rows_per_lane": 5}, {"rows_per_lane": -1}, {"fallback": 2},
                 {"fallback": -1}):
        st, ipiv, jpiv, inf, rep, _ = dlu.getrf_device(4, 4, 1234, 4, options=opts)
        assert st == dlu.Status.INVALID_ARGUMENT and inf == -1, opts


def test_column_variant_argument_rules():
    # delayed_lu/delayed_dgetrf_column.cpp:101-112: the same argument rules as delayed_getrf
    ip = np.zeros(4, dtype=np.int32)
    L = dlu.lib()
    for m, n, a, lda in ((-1, 3, 1234, 3), (3, 3, 1234, 2), (3, 3, None, 3)):
        info = ctypes.c_int(7)
        for fn in (L.dlu_getrf_column_device, L.dlu_getrf_column_host):
            assert fn(m, n, a, lda, ip.ctypes.data_as(dlu._ip), ip.ctypes.data_as(dlu._ip),
                      ctypes.byref(info), None) == dlu.Status.INVALID_ARGUMENT
            assert info.value == -1
    info = ctypes.c_int(7)
    assert L.dlu_getrf_column_device(0, 5, None, 1, ip.ctypes.data_as(dlu._ip),
                                     ip.ctypes.data_as(dlu._ip), ctypes.byref(info), None) == 0
    assert info.value == 0


@pytest.mark.parametrize("m,n", [(0, 5), (5, 0), (0, 0)])
def test_empty_problems(m, n):
    # delayed_dgetrf_cuda.cu:411-419 (device: info = min(m,n) = 0) and :456-460 (host: info = 0)
    ip = np.zeros(1, dtype=np.int32)
    info = ctypes.c_int(7)
    L = dlu.lib()
    assert L.dlu_getrf_device(m, n, None, max(1, m), ip.ctypes.data_as(dlu._ip),
                              ip.ctypes.data_as(dlu._ip), ctypes.byref(info), None) == 0
    assert info.value == 0
    info.value = 7
    assert L.dlu_getrf_host(m, n, None, max(1, m), ip.ctypes.data_as(dlu._ip),
                            ip.ctypes.data_as(dlu._ip), ctypes.byref(info), None) == 0
    assert info.value == 0


def test_gemm_entry_rejects_bad_shapes():
    assert dlu.gemm_sub_device(-1, 2, 2, 1, 1, 1, 4) == dlu.Status.INVALID_ARGUMENT
    assert dlu.gemm_sub_device(8, 2, 2, 1, 1, 1, 4) == dlu.Status.INVALID_ARGUMENT


def test_tuned_table_is_what_the_library_uses():
    """The autotuned table (csrc/dlu_tuning.inc, written by autotune.py) is compiled in: the
    library's choices equal the table's entries for each depth and size class."""
    import os
    import re
    text = open(os.path.join(os.path.dirname(dlu.LIB_PATH), "..", "csrc", "dlu_tuning.inc")).read()
    rows = re.search(r"kTunedGemmTiles\[2\]\[5\] = \{(.*)\};", text).group(1)
    table = [[int(x) for x in re.findall(r"\d+", r)] for r in re.findall(r"\{([^{}]*)\}", rows)]
    cutoff = int(re.search(r"kTunedBatchCutoff = (\d+);", text).group(1))
    assert len(table) == 2 and all(len(r) == 5 and all(0 <= c < 4 for c in r) for r in table)
    L = dlu.lib()
    sizes = [(896, 896), (1920, 1920), (2816, 2816), (5632, 5632), (16128, 16128)]
    for d, K in enumerate((128, 256)):
        for s, (M, N) in enumerate(sizes):
            assert L.dlu_tuned_gemm_config(M, N, K) == table[d][s], (K, M)
    assert L.dlu_tuned_batch_cutoff() == min(cutoff, 128) and 32 <= cutoff <= 128
    assert L.dlu_tuned_gemm_config(0, 5, 5) == -1
    assert dlu.gemm_sub_device(8, 8, 8, 1, 1, 1, 8, config=4) == dlu.Status.INVALID_ARGUMENT


def test_python_struct_layouts_match_the_header(tmp_path):
    """Every ctypes mirror of a dlu.h struct has the C struct's size and field offsets (gcc on
    the header itself)."""
    import os
    mirrors = {"dlu_options": dlu.Options, "dlu_report": dlu.Report,
               "dlu_mixed_options": dlu.MixedOptions, "dlu_mixed_report": dlu.MixedReport,
               "dlu_batch_item": dlu.BatchItem, "dlu_server_config": dlu.ServerConfig,
               "dlu_request": dlu.Request, "dlu_completion": dlu.Completion}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dlu.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    include = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", include, str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.splitlines():
        c, f, v = line.split()
        got[(c, f)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
