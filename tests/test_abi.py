"""The C-ABI library loads (no GPU needed) and exports every function the
headers in include/ declare."""
import ctypes
import os
import re

from conftest import ROOT
from hethyppar_amd import hbs

DECL = re.compile(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", re.M)


def declared_functions():
    names = {}
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"^\s*#[^\n]*", "", text, flags=re.M)
        text = re.sub(r"typedef[^;]*;", "", text)
        for m in DECL.finditer(text):
            names.setdefault(m.group(1), f)
    return names


def test_headers_declare_something():
    names = declared_functions()
    assert "hbs_bicgstab_f64" in names and "hbs_shard_spmv" in names


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(hbs.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert missing == []


def test_error_path_without_device():
    # argument errors come back as codes with a message, never an exit
    rc = hbs.lib().hbs_csr_destroy(None)
    assert rc == 0
    rc = hbs.lib().hbs_set_spmv_plan(100, 1)
    assert rc == -1 and b"block_nnz" in hbs.lib().hbs_last_error()


def test_plan_knob_validation_without_device():
    L = hbs.lib()
    assert L.hbs_set_spmv_plan(16384 + 256, 1) == -1 and b"block_nnz" in L.hbs_last_error()
    assert L.hbs_set_spmv_plan(16384, 64) == 0
    assert L.hbs_set_spmv_pass(6) == -1 and b"entries_per_thread" in L.hbs_last_error()
    assert L.hbs_set_spmv_pass(8) == 0 and L.hbs_set_spmv_pass(0) == 0
    assert L.hbs_set_spmv_groups(-1) == -1 and b"max_groups" in L.hbs_last_error()
    assert L.hbs_set_spmv_groups(0) == 0
    assert L.hbs_set_spmv_plan(0, 64) == 0
    assert L.hbs_set_spmv_xseg(3) == -1 and b"mode" in L.hbs_last_error()
    assert L.hbs_set_spmv_xseg(2) == 0


def test_spmm_and_block_knob_validation_without_device():
    L = hbs.lib()
    assert L.hbs_set_spmm_order(2) == -1 and b"mode" in L.hbs_last_error()
    assert L.hbs_set_spmm_order(1) == 0
    assert L.hbs_set_spmm_tuning(0, 12, 16, 1) == -1 and b"groups" in L.hbs_last_error()
    assert L.hbs_set_spmm_tuning(0, 0, 17, 1) == -1
    assert L.hbs_set_spmm_tuning(0, 0, 16, 2) == -1
    assert L.hbs_set_spmm_tuning(0, 0, 16, 1) == 0
    assert L.hbs_set_block_tuning(48, 0) == -1 and b"tile" in L.hbs_last_error()
    assert L.hbs_set_block_tuning(64, 3) == -1
    assert L.hbs_set_block_tuning(0, 0) == 0
    nr, ln = ctypes.c_int(), ctypes.c_int()
    assert L.hbs_block_tuning(None, 4, ctypes.byref(nr), ctypes.byref(ln), ctypes.byref(ln)) == -1


def test_wait_limit_knob_validation_without_device():
    L = hbs.lib()
    assert L.hbs_set_wait_limit(ctypes.c_double(0.0)) == -1 and b"wait limit" in L.hbs_last_error()
    assert L.hbs_set_wait_limit(ctypes.c_double(-1.0)) == -1
    assert L.hbs_set_wait_limit(ctypes.c_double(600.0)) == 0


def test_csr_create_rejects_columns_past_32bit_x_offsets():
    """The SpMV gathers x through 32-bit byte offsets, so a matrix with
    n * 8 > 2^32 - 1 columns is refused before anything touches a device.
    (The context handle is never dereferenced on this path: argument checks
    come first.)"""
    L = hbs.lib()
    rp = (ctypes.c_int * 2)(0, 0)
    out = ctypes.c_void_p()
    fake_ctx = ctypes.c_void_p(1)
    for n, ok in ((1 << 29, False), ((1 << 29) + 5, False)):
        rc = L.hbs_csr_create(fake_ctx, 1, n, 0, rp, None, None, 0, ctypes.byref(out))
        assert rc == -1 and b"32-bit" in L.hbs_last_error()
    rc = L.hbs_csr_create(fake_ctx, -1, 4, 0, rp, None, None, 0, ctypes.byref(out))
    assert rc == -1


def test_plain_c_caller_compiles_against_headers(tmp_path):
    """A C11 program written against include/hhp_hip.h and hhp_dp_kernels.h
    (the reference harness' call sequence) compiles warning-free and links."""
    from c_caller import build_caller
    exe = build_caller(tmp_path)
    assert os.path.exists(exe)
