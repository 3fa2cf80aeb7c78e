"""The C-ABI library is built, loads, and exports every function include/itcpd.h declares.
No compute call is made (no GPU is needed)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "itcpd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(itcpd_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_path():
    fns = declared_functions()
    for need in ("itcpd_plan_create", "itcpd_run", "itcpd_set_cpd", "itcpd_get_cpd",
                 "itcpd_random_cpd", "itcpd_mttkrp", "itcpd_solve_ls", "itcpd_fit",
                 "itcpd_compute_krp", "itcpd_row_norm_post_solve", "itcpd_last_error"):
        assert need in fns


def test_library_exports_every_declared_symbol(itcpd):
    lib = itcpd._lib.load()
    for fn in declared_functions():
        assert hasattr(lib, fn), fn


def test_ctypes_table_covers_header(itcpd):
    assert set(itcpd._lib.SIGNATURES) == set(declared_functions())


def test_argument_errors_need_no_device(itcpd):
    lib = itcpd._lib.load()
    h = ctypes.c_void_p()
    dims = (ctypes.c_int64 * 1)(5)
    rc = lib.itcpd_plan_create(ctypes.byref(h), 1, dims, 3, 8, ctypes.c_void_p(1), 0, 0, 0, 0)
    assert rc == itcpd._lib.ITCPD_E_ARG
    assert b"order" in lib.itcpd_last_error()
    rc = lib.itcpd_plan_create(ctypes.byref(h), 3, (ctypes.c_int64 * 3)(4, 4, 4), 0, 8,
                               ctypes.c_void_p(1), 0, 0, 0, 0)
    assert rc == itcpd._lib.ITCPD_E_ARG and b"rank" in lib.itcpd_last_error()
    assert lib.itcpd_version().startswith(b"itcpd-hip")


def test_structs_match_header_layout(itcpd):
    L = itcpd._lib
    assert ctypes.sizeof(L.CheckStruct) == 8 * 4 + 4 * 8
    assert ctypes.sizeof(L.KernelStats) == 32
