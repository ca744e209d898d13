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
    assert L.rg_abi_version() == 2
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


def _struct_fields(name):
    src = open(os.path.join(REPO, "include", "rg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, flags=re.S).group(1)
    return [m.group(2) for m in re.finditer(r"\b([a-z0-9_]+)\s+\**\s*([a-z0-9_]+)\s*;", body)]


@pytest.mark.parametrize("cname,pyname", [("rg_grid_stats", "GridStats"),
                                          ("rg_grid_params", "GridParams"),
                                          ("rg_flat_vis", "FlatVis"), ("rg_wkern", "WKern")])
def test_struct_layouts_match_the_header(cname, pyname):
    """The ctypes structures mirror the header's field order (a mismatch would
    shift every later field)."""
    ours = [f[0] for f in getattr(rgamd, pyname)._fields_]
    assert _struct_fields(cname) == ours


def _dummy_call(n, mode=1, subgrid=0):
    """rg_grid_noalloc with host buffers standing in for device arrays: only
    calls that the host-side checks reject are made (nothing is read)."""
    buf = np.zeros(64, np.complex128)
    ptr = buf.ctypes.data
    fv = rgamd.FlatVis(ptr, ptr, ptr, ptr, n, rgamd.RG_C128, 0)
    wk = rgamd.WKern(ptr, rgamd.RG_C128, 1, 8, 16, -5.0, 1.0)
    p = rgamd.params(2048, 0.1, mode=mode, subgrid=subgrid, margin=16 if subgrid else 0,
                     winc=10.0 if subgrid else 0.0)
    return rgamd.lib().rg_grid_noalloc(C.byref(p), C.byref(fv), C.byref(wk), ptr, rgamd.RG_C128,
                                       0, None), buf


def test_noalloc_rejects_streams_beyond_its_scratch():
    code, _ = _dummy_call(50_000_000)
    assert code == rgamd.RG_ERR_PARAM
    msg = rgamd.lib().rg_last_error().decode()
    assert "static scratch" in msg and "accumulate" in msg


def test_noalloc_is_batched_wprojection_only():
    assert _dummy_call(10, mode=rgamd.RG_MODE_PLAIN)[0] == rgamd.RG_ERR_UNSUPPORTED
    assert _dummy_call(10, subgrid=512)[0] == rgamd.RG_ERR_UNSUPPORTED
