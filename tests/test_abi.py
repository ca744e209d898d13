"""The C ABI: every function declared in include/rg.h is exported by librg.so
and bound by the ctypes layer; host-only entry points reject bad arguments
with error codes (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import rgamd

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(REPO, "include", "rg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rg_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported_and_bound():
    names = declared()
    assert len(names) >= 30
    lib = C.CDLL(rgamd.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    bound = {s[0] for s in rgamd.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_version_and_error_codes():
    L = rgamd.lib()
    assert L.rg_abi_version() == 1
    out = np.empty(16, np.complex128)
    assert L.rg_wkern_generate(out.ctypes.data, 7, 1, 2, 2, 0.1, 0.0, 1.0, 0, 0.0, 1) == rgamd.RG_ERR_PARAM
    assert b"type" in L.rg_last_error()
    assert L.rg_wkern_generate(None, 0, 1, 2, 2, 0.1, 0.0, 1.0, 0, 0.0, 1) == rgamd.RG_ERR_PARAM
    assert L.rg_pswf_eval(1.0, None, -1, None) == rgamd.RG_ERR_PARAM
    assert L.rg_set_file_mode(5) == rgamd.RG_ERR_PARAM


def test_null_context_rejected():
    L = rgamd.lib()
    p = rgamd.params(64, 0.1)
    assert L.rg_grid(None, C.byref(p), None, None, None, 0, None) == rgamd.RG_ERR_PARAM
    assert L.rg_context_destroy(None) == rgamd.RG_OK
