"""The C-ABI library loads without a GPU and exports every entry point its
header declares; the ctypes layer binds each of them.  No compute call is
made here."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "dqmc_gpu.h")
LIB = os.path.join(ROOT, "sai-qmc_amd", "libdqmc_gpu.so")


def declared():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    names = re.findall(r"\b(dqmc_gpu_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    assert len(names) >= 20
    for must in ("dqmc_gpu_create", "dqmc_gpu_sweep", "dqmc_gpu_read_bins", "dqmc_gpu_udv_inv_one_plus",
                 "dqmc_gpu_destroy", "dqmc_gpu_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libdqmc_gpu.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_layer_binds_every_symbol(eng_mod):
    names = declared()
    bound = set(eng_mod.SIGNATURES)
    assert set(names) == bound, (set(names) ^ bound)


def test_last_error_without_gpu():
    lib = C.CDLL(LIB)
    lib.dqmc_gpu_last_error.restype = C.c_char_p
    assert isinstance(lib.dqmc_gpu_last_error(), bytes)
