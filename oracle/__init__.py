"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``liboracle.so``, the CPU restatement of FTK2's CPU engine
(include/ftk2/core/engine.hpp execute()/get_complex() and the predicates it runs).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
may import this package, and only as the checker (or the timed CPU baseline).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

PRED_CP, PRED_CONTOUR, PRED_FIBER = 0, 1, 2
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P, U64P, I64P, DP = C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_int64), C.POINTER(C.c_double)
        L.oracle_run.restype = P
        L.oracle_run.argtypes = [C.c_int, U64P, U64P, U64P, C.c_int, C.c_int, C.c_int, C.c_double, DP,
                                 C.c_int, C.POINTER(C.c_void_p), I64P, C.c_void_p, I64P, C.c_int, C.c_uint64]
        L.oracle_run_stream.restype = P
        L.oracle_run_stream.argtypes = [C.c_int, U64P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, DP,
                                        C.c_int, C.POINTER(C.c_void_p), C.c_int]
        L.oracle_error.restype = C.c_char_p
        L.oracle_n_nodes.restype = C.c_uint64
        L.oracle_n_nodes.argtypes = [P]
        L.oracle_node_dim.argtypes = [P]
        L.oracle_nodes.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_n_conn.argtypes = [P]
        L.oracle_conn_dim.argtypes = [P, C.c_int]
        L.oracle_conn_size.restype = C.c_uint64
        L.oracle_conn_size.argtypes = [P, C.c_int]
        L.oracle_conn.argtypes = [P, C.c_int, C.c_void_p]
        L.oracle_n_components.restype = C.c_uint64
        L.oracle_n_components.argtypes = [P]
        L.oracle_free.argtypes = [P]
        L.oracle_lean_run.restype = P
        L.oracle_lean_run.argtypes = [C.c_int, U64P, U64P, U64P, C.c_int, C.c_int, C.c_int, C.c_double, DP,
                                      C.c_int, C.POINTER(C.c_void_p), I64P, C.c_int]
        L.oracle_lean_sizes.argtypes = [P, U64P, U64P, U64P]
        L.oracle_lean_nodes.argtypes = [P, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_lean_edges.argtypes = [P, C.c_uint64, C.c_uint64, C.c_void_p]
        L.oracle_lean_free.argtypes = [P]
        L.oracle_sign.argtypes = [C.c_int, C.c_double, C.c_uint64, C.c_double]
        L.oracle_det2.argtypes = [C.c_int, DP, DP, C.c_uint64, C.c_uint64, C.c_double]
        L.oracle_det3.argtypes = [C.c_int, DP, DP, DP, C.c_double]
        L.oracle_det3_level0.argtypes = [I64P, I64P, I64P, I64P, U64P]
        L.oracle_origin_inside.argtypes = [C.c_int, C.c_int, DP, U64P, C.c_double]
        L.oracle_solve.argtypes = [C.c_int, C.c_int, DP, DP]
        L.oracle_origin_inside_many.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_void_p, C.c_void_p, C.c_double,
                                                C.c_void_p]
        L.oracle_chains.argtypes = [C.c_int, C.c_int, C.c_void_p]
        L.oracle_encode_simplex_id.restype = C.c_uint64
        L.oracle_encode_simplex_id.argtypes = [C.c_int, U64P, C.c_int, U64P]
        _lib = L
    return _lib


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint64))


def _dp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _collect(h):
    L = lib()
    if not h:
        raise RuntimeError("oracle failed: " + L.oracle_error().decode())
    try:
        n = L.oracle_n_nodes(h)
        nv = L.oracle_node_dim(h) + 1
        verts = np.zeros((n, nv), np.uint64)
        lam = np.zeros((n, nv), np.float64)
        bary = np.zeros((n, nv), np.float32)
        scal = np.zeros(n, np.float32)
        track = np.zeros(n, np.uint64)
        L.oracle_nodes(h, verts.ctypes.data, lam.ctypes.data, bary.ctypes.data, scal.ctypes.data,
                       track.ctypes.data)
        conn = {}
        for i in range(L.oracle_n_conn(h)):
            dim = L.oracle_conn_dim(h, i)
            sz = L.oracle_conn_size(h, i)
            buf = np.zeros(sz, np.uint32)
            L.oracle_conn(h, i, buf.ctypes.data)
            conn[dim] = buf.reshape(-1, dim + 1)
        return {"verts": verts, "lam": lam, "bary": bary, "scalar": scal, "track": track,
                "conn": conn, "n_components": int(L.oracle_n_components(h))}
    finally:
        L.oracle_free(h)


def _dtype_code(arrays):
    """0 for float32, 1 for float64; every array must share it (the C side reads all
    of them with the first one's element size)."""
    dt = arrays[0].dtype
    if dt not in (np.float32, np.float64):
        raise TypeError(f"field arrays must be float32 or float64, not {dt}")
    for a in arrays:
        if a.dtype != dt:
            raise TypeError(f"field arrays mix {dt} and {a.dtype}")
    return 0 if dt == np.float32 else 1


def run(comps, kind, m, dims=None, offset=None, gdims=None, thr=(0.0, 0.0), Q=1e6, scalar=None,
        nthreads=0, time_stride=0):
    """Run the CPU engine on a regular mesh.

    comps: list of numpy arrays (float32 or float64), indexed [x, y, (z), t] on local
    coordinates, any strides (the reference's legacy separate-array mode).
    time_stride: nonzero for an ExtrudedSimplicialMesh over a regular base (vertex
    id = spatial id + t * time_stride, see extruded_time_stride).
    """
    comps = list(comps)
    dtype = _dtype_code(comps)
    d = comps[0].ndim
    if dims is None:
        dims = comps[0].shape
    item = comps[0].itemsize
    ptrs = (C.c_void_p * len(comps))(*[c.ctypes.data for c in comps])
    st = np.zeros(len(comps) * 4, np.int64)
    for i, c in enumerate(comps):
        assert c.dtype == comps[0].dtype
        for j in range(d):
            st[i * 4 + j] = c.strides[j] // item
    dims_a, dims_p = _u64(list(dims) + [1] * (4 - d))
    off_a, off_p = _u64((list(offset) if offset is not None else [0] * d) + [0] * (4 - d))
    g_a, g_p = _u64((list(gdims) if gdims is not None else list(dims)) + [1] * (4 - d))
    thr_a, thr_p = _dp(list(thr) + [0.0] * (2 - len(thr)))
    sst = np.zeros(4, np.int64)
    sptr = None
    if scalar is not None:
        assert scalar.dtype == comps[0].dtype
        sptr = scalar.ctypes.data
        for j in range(d):
            sst[j] = scalar.strides[j] // item
    h = lib().oracle_run(d, dims_p, off_p, g_p, kind, m, dtype, Q, thr_p, len(comps), ptrs,
                         st.ctypes.data_as(C.POINTER(C.c_int64)), sptr,
                         sst.ctypes.data_as(C.POINTER(C.c_int64)), nthreads, C.c_uint64(time_stride))
    return _collect(h)


class Lean:
    """Full-size mode (lean.hpp) for star-edge predicates: the same FeatureComplex as
    run(), held compactly; read it in ranges with nodes() / edges()."""

    def __init__(self, comps, kind, m, dims=None, offset=None, gdims=None, thr=(0.0, 0.0), Q=1e6, nthreads=0):
        comps = list(comps)
        dtype = _dtype_code(comps)
        d = comps[0].ndim
        dims = comps[0].shape if dims is None else dims
        item = comps[0].itemsize
        ptrs = (C.c_void_p * len(comps))(*[c.ctypes.data for c in comps])
        st = np.zeros(len(comps) * 4, np.int64)
        for i, c in enumerate(comps):
            assert c.dtype == comps[0].dtype
            for j in range(d):
                st[i * 4 + j] = c.strides[j] // item
        dims_a, dims_p = _u64(list(dims) + [1] * (4 - d))
        off_a, off_p = _u64((list(offset) if offset is not None else [0] * d) + [0] * (4 - d))
        g_a, g_p = _u64((list(gdims) if gdims is not None else list(dims)) + [1] * (4 - d))
        thr_a, thr_p = _dp(list(thr) + [0.0] * (2 - len(thr)))
        L = lib()
        self._h = L.oracle_lean_run(d, dims_p, off_p, g_p, kind, m, dtype, Q, thr_p, len(comps), ptrs,
                                    st.ctypes.data_as(C.POINTER(C.c_int64)), nthreads)
        if not self._h:
            raise RuntimeError("oracle failed: " + L.oracle_error().decode())
        self.m = m
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        L.oracle_lean_sizes(self._h, C.byref(a), C.byref(b), C.byref(c))
        self.n_nodes, self.n_edges, self.n_components = a.value, b.value, c.value

    def nodes(self, begin=0, count=None):
        count = self.n_nodes - begin if count is None else count
        nv = self.m + 1
        verts = np.zeros((count, nv), np.uint64)
        lam = np.zeros((count, nv), np.float64)
        track = np.zeros(count, np.uint64)
        lib().oracle_lean_nodes(self._h, begin, count, verts.ctypes.data, lam.ctypes.data, track.ctypes.data)
        return verts, lam, track

    def edges(self, begin=0, count=None):
        count = self.n_edges - begin if count is None else count
        out = np.zeros((count, 2), np.uint32)
        lib().oracle_lean_edges(self._h, begin, count, out.ctypes.data)
        return out

    def close(self):
        if self._h:
            lib().oracle_lean_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def extruded_time_stride(spatial_global_dims):
    """ExtrudedSimplicialMesh::n_spatial_verts_ for a regular base
    (mesh.hpp:547-552): grid_index_to_id(global dims) = sum_j g_j * prod_{i<j} g_i."""
    r, mul = 0, 1
    for g in spatial_global_dims:
        r += int(g) * mul
        mul *= int(g)
    return r


def run_stream(steps, kind, m, thr=(0.0, 0.0), Q=1e6, nthreads=0):
    """execute_stream(): steps[t] is a list of per-component spatial arrays (x fastest)."""
    nt = len(steps)
    ncomp = len(steps[0])
    sp = steps[0][0].shape
    dtype = _dtype_code([a for st in steps for a in st])
    flat = []
    for t in range(nt):
        for c in range(ncomp):
            a = steps[t][c]
            assert a.flags.f_contiguous and a.shape == sp
            flat.append(a)
    ptrs = (C.c_void_p * len(flat))(*[a.ctypes.data for a in flat])
    sp_a, sp_p = _u64(list(sp))
    thr_a, thr_p = _dp(list(thr) + [0.0] * (2 - len(thr)))
    h = lib().oracle_run_stream(len(sp), sp_p, nt, kind, m, dtype, Q, thr_p, ncomp, ptrs, nthreads)
    return _collect(h)


# ---- unit-level helpers -------------------------------------------------------
def sign(v, idx, Q=1e6, dtype=1):
    return lib().oracle_sign(dtype, float(v), int(idx), Q)


def det2(a, b, i0, i1, Q=1e6, dtype=1):
    a_, ap = _dp(a)
    b_, bp = _dp(b)
    return lib().oracle_det2(dtype, ap, bp, int(i0), int(i1), Q)


def det3(a, b, c, Q=1e6, dtype=1):
    a_, ap = _dp(a)
    b_, bp = _dp(b)
    c_, cp = _dp(c)
    return lib().oracle_det3(dtype, ap, bp, cp, Q)


def det3_level0(qa, qb, qc):
    arrs = [np.ascontiguousarray(x, dtype=np.int64) for x in (qa, qb, qc)]
    hi = C.c_int64()
    lo = C.c_uint64()
    P = C.POINTER(C.c_int64)
    lib().oracle_det3_level0(arrs[0].ctypes.data_as(P), arrs[1].ctypes.data_as(P),
                             arrs[2].ctypes.data_as(P), C.byref(hi), C.byref(lo))
    return (hi.value << 64) + lo.value


def origin_inside(values, idx, Q=1e6, dtype=1):
    v = np.asarray(values, dtype=np.float64)
    K = v.shape[1]
    v_, vp = _dp(v.ravel())
    i_, ip = _u64(idx)
    return bool(lib().oracle_origin_inside(K, dtype, vp, ip, Q))


def origin_inside_many(values, idx, Q=1e6, dtype=1):
    """values (n, K+1, K), idx (n, K+1) -> bool (n,)"""
    v = np.ascontiguousarray(values, dtype=np.float64)
    n, kp1, K = v.shape
    ix = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.zeros(n, np.int32)
    lib().oracle_origin_inside_many(K, dtype, n, v.ctypes.data, ix.ctypes.data, Q, out.ctypes.data)
    return out.astype(bool)


def solve(values, dtype=1):
    v = np.asarray(values, dtype=np.float64)
    K = v.shape[1]
    v_, vp = _dp(v.ravel())
    lam = np.zeros(K + 1, np.float64)
    ok = lib().oracle_solve(K, dtype, vp, lam.ctypes.data_as(C.POINTER(C.c_double)))
    return bool(ok), lam


def chains(d, m):
    out = np.zeros(64 * 5, np.uint8)
    n = lib().oracle_chains(d, m, out.ctypes.data)
    return out[: n * (m + 1)].reshape(n, m + 1)


def encode_simplex_id(gdims, verts):
    d = len(gdims)
    g_a, g_p = _u64(list(gdims))
    v_a, v_p = _u64(list(verts))
    return int(lib().oracle_encode_simplex_id(d, g_p, len(verts) - 1, v_p))


def positions(verts, bary, gdims, time_stride=0):
    """Point positions and "Time" as the reference's VTK writer computes them
    (utils/vtk.hpp:26-59, complex_to_vtk_points): per node, phys_pos[j] +=
    (double)barycentric_coords[0][i] * coordinate j of vertex i, for the node's
    vertices in order, over every mesh dimension; the point is phys_pos[0..2] stored as
    float (vtkPoints), Time = (float) phys_pos[last] when the mesh has a time axis beyond
    its spatial dimension (an ExtrudedSimplicialMesh), else 0.  Vertex coordinates are
    the global grid indices (mesh.hpp:420-451 id_to_grid_index; extruded: layer =
    id / n_spatial_verts_, mesh.hpp:711-717).  Returns ((n, 3) float32, (n,) float32)."""
    verts = np.asarray(verts, dtype=np.uint64)
    n, nv = verts.shape
    d = len(gdims)
    coords = np.zeros((n, nv, d), np.float64)
    rest = verts.copy()
    sd = d
    if time_stride:
        coords[:, :, d - 1] = (rest // np.uint64(time_stride)).astype(np.float64)
        rest = rest % np.uint64(time_stride)
        sd = d - 1
    for j in range(sd):
        coords[:, :, j] = (rest % np.uint64(gdims[j])).astype(np.float64)
        rest = rest // np.uint64(gdims[j])
    phys = np.zeros((n, d), np.float64)
    for i in range(nv):
        b = np.asarray(bary[:, i], dtype=np.float32).astype(np.float64)
        for j in range(d):
            phys[:, j] = phys[:, j] + b * coords[:, i, j]
    pts = np.zeros((n, 3), np.float32)
    for j in range(min(3, d)):
        pts[:, j] = phys[:, j].astype(np.float32)
    time = phys[:, d - 1].astype(np.float32) if time_stride else np.zeros(n, np.float32)
    return pts, time


def attributes(verts, bary, specs, gdims, offset, dtype):
    """Attribute values of FeatureElement::attributes (engine.hpp:1144-1201).

    verts: (n, m+1) ascending global vertex ids of each node; bary: (n, m+1) float32
    barycentric coordinates; specs: [(slot, [arrays indexed by local coordinates])]
    in order (one array: its value; several: the Euclidean norm, the reference's
    "magnitude"); gdims/offset: global extents and local offset of the mesh;
    dtype: the predicate dtype, in which each sum sum_i bary_i * value_i runs term by
    term from i = 0.  Returns (n, 16) float32, zero where no spec writes."""
    T = np.dtype(dtype).type
    verts = np.asarray(verts, dtype=np.uint64)
    n, nv = verts.shape
    d = len(gdims)
    out = np.zeros((n, 16), np.float32)
    # local coordinates of every vertex
    coords = []
    rest = verts.copy()
    for j in range(d):
        coords.append((rest % np.uint64(gdims[j])).astype(np.int64) - int(offset[j]))
        rest = rest // np.uint64(gdims[j])
    for slot, arrays in specs:
        acc = np.zeros(n, dtype=T)
        for k in range(nv):
            idx = tuple(c[:, k] for c in coords)
            if len(arrays) == 1:
                v = np.asarray(arrays[0])[idx].astype(T)
            else:
                msq = np.zeros(n, dtype=T)
                for a in arrays:
                    x = np.asarray(a)[idx].astype(T)
                    msq = (msq + x * x).astype(T)
                v = np.sqrt(msq).astype(T)
            acc = (acc + bary[:, k].astype(T) * v).astype(T)
        out[:, slot] = acc.astype(np.float32)
    return out
