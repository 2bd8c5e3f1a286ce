"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``_ref/libftk2ref.so``: the reference's own sos.hpp and mesh.hpp
(hguo/ftk2 include/ftk2/core/) compiled by ``make -C oracle ref`` from the reference
tree.  Only tests use it, to check the oracle's restatement; it exists only where the
reference tree is present (``available()``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_ref", "libftk2ref.so")
REF_TREE = os.environ.get("FTK2_REFERENCE", "/root/reference")
_lib = None


def build() -> bool:
    """Compile the reference build when the reference tree is present."""
    if not os.path.isdir(os.path.join(REF_TREE, "include", "ftk2", "core")):
        return False
    subprocess.run(["make", "-s", "-C", _HERE, "ref", f"REF={REF_TREE}"], check=True, stdout=subprocess.DEVNULL)
    return True


def available() -> bool:
    return os.path.exists(_LIB)


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(_LIB)
        U64P, DP = C.POINTER(C.c_uint64), C.POINTER(C.c_double)
        L.ref_sign.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_double]
        L.ref_det2.argtypes = [C.c_int, DP, DP, C.c_uint64, C.c_uint64, C.c_double]
        L.ref_det3.argtypes = [C.c_int, DP, DP, DP, C.c_double]
        L.ref_origin_inside_many.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_double,
                                             C.c_void_p]
        L.ref_num_simplex_types.argtypes = [C.c_int, C.c_int]
        L.ref_lut.argtypes = [C.c_int] * 5
        L.ref_iterate_simplices.restype = C.c_uint64
        L.ref_iterate_simplices.argtypes = [C.c_int, U64P, U64P, U64P, C.c_int, C.c_void_p, C.c_uint64]
        L.ref_encode_simplex_id.restype = C.c_uint64
        L.ref_encode_simplex_id.argtypes = [C.c_int, U64P, C.c_int, U64P]
        L.ref_coface_neighbors.argtypes = [U64P, U64P, U64P, U64P, C.c_void_p, C.c_int]
        _lib = L
    return _lib


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint64))


def _dp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def sign(v, idx, Q=1e6, dtype=1):
    return lib().ref_sign(dtype, float(v), int(idx), Q)


def det2(a, b, i0, i1, Q=1e6, dtype=1):
    a_, ap = _dp(a)
    b_, bp = _dp(b)
    return lib().ref_det2(dtype, ap, bp, int(i0), int(i1), Q)


def det3(a, b, c, Q=1e6, dtype=1):
    a_, ap = _dp(a)
    b_, bp = _dp(b)
    c_, cp = _dp(c)
    return lib().ref_det3(dtype, ap, bp, cp, Q)


def origin_inside_many(values, idx, Q=1e6, dtype=1):
    """values (n, K+1, K), idx (n, K+1) -> bool (n,)"""
    v = np.ascontiguousarray(values, dtype=np.float64)
    n, kp1, K = v.shape
    ix = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.zeros(n, np.int32)
    lib().ref_origin_inside_many(K, dtype, n, v.ctypes.data, ix.ctypes.data, Q, out.ctypes.data)
    return out.astype(bool)


def lut(d, m):
    """Rows of cpu_lut_D_M: (types, m+1, d) vertex offsets."""
    n = lib().ref_num_simplex_types(d, m)
    return np.array([[[lib().ref_lut(d, m, t, vi, j) for j in range(d)] for vi in range(m + 1)] for t in range(n)],
                    dtype=np.int64).reshape(n, m + 1, d)


def iterate_simplices(dims, k, offset=None, gdims=None):
    d = len(dims)
    dims_a, dp = _u64(dims)
    off_a, op = _u64(offset) if offset is not None else (None, None)
    g_a, gp = _u64(gdims) if gdims is not None else (None, None)
    n = lib().ref_iterate_simplices(d, dp, op, gp, k, None, 0)
    out = np.zeros((n, k + 1), np.uint64)
    lib().ref_iterate_simplices(d, dp, op, gp, k, out.ctypes.data, n)
    return out


def encode_simplex_id(gdims, verts):
    g_a, gp = _u64(gdims)
    v_a, vp = _u64(verts)
    return int(lib().ref_encode_simplex_id(len(gdims), gp, len(verts) - 1, vp))


def coface_neighbors(dims, tet, offset=None, gdims=None):
    dims_a, dp = _u64(dims)
    off_a, op = _u64(offset) if offset is not None else (None, None)
    g_a, gp = _u64(gdims) if gdims is not None else (None, None)
    t_a, tp = _u64(tet)
    out = np.zeros((256, 4), np.uint64)
    n = lib().ref_coface_neighbors(dp, op, gp, tp, out.ctypes.data, 256)
    assert n <= 256
    return out[:n]
