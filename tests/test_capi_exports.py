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
    L = _lib.lib()
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


def test_one_hip_runtime_when_torch_comes_later():
    """Loading the library and importing torch afterwards leaves one HIP runtime and one
    HSA runtime mapped (two of them make torch's device initialisation fail)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from ftk2_amd import _lib\n"
            "_lib.lib()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "print(len({l.split()[-1] for l in maps if 'libamdhip64' in l}),"
            " len({l.split()[-1] for l in maps if 'libhsa-runtime64' in l}))\n" % ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.split() == ["1", "1"], out.stdout
