"""The C-ABI library loads and exports every entry point include/myqc_hip.h
declares (no compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HDR = os.path.join(ROOT, "include", "myqc_hip.h")
LIB = os.path.join(ROOT, "myqc_amd", "libmyqc_hip.so")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(myqc_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for must in ["myqc_init", "myqc_finalize", "myqc_set_system", "myqc_eri_full", "myqc_eri_build", "myqc_fock_G",
                 "myqc_density_rhf", "myqc_ao2mo_ovov", "myqc_mp2_rhf", "myqc_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libmyqc_hip.so not built")
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol():
    from myqc_amd.hip import EXPORTS
    assert sorted(EXPORTS) == declared()


def test_null_context_is_rejected_without_gpu():
    L = ctypes.CDLL(LIB)
    L.myqc_finalize.argtypes = [ctypes.c_void_p]
    L.myqc_last_error.argtypes = [ctypes.c_void_p]
    L.myqc_last_error.restype = ctypes.c_char_p
    assert L.myqc_finalize(None) == -1
    assert L.myqc_last_error(None) == b"null context"
