"""The C-ABI library loads on a machine without a GPU and exports every entry
point that include/ftk2h.h declares (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ftk2h.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ftk2h_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from ftk2_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert "ftk2h_extract_regular" in syms
    for s in syms:
        assert hasattr(L, s), s
    # the Python binding knows every entry point
    assert set(syms) <= set(_lib.SIGNATURES)


def test_no_gpu_reports_error_not_crash():
    import torch
    if torch.cuda.is_available():
        return
    from ftk2_amd import _lib
    p = ctypes.c_void_p()
    rc = _lib.lib().ftk2h_ctx_create(0, ctypes.byref(p))
    assert rc != 0
    assert _lib.lib().ftk2h_last_error()
