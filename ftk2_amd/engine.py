"""Python mirror of FTK2's SimplicialEngine interface over the C ABI (libftk2h.so).

Mirrors, on a RegularSimplicialMesh:
  RegularSimplicialMesh(dims, offset, global_dims)   include/ftk2/core/mesh.hpp:256-265
  CriticalPointPredicate / ContourPredicate / FiberPredicate   predicate.hpp:49,135,340
  SimplicialEngine(mesh, pred).execute(data, var_names) / get_complex()
                                                     engine.hpp:84,96,822
  FeatureComplex (vertices sorted by simplex, connectivity sorted)  complex.hpp:14-58

Arrays are indexed like ftk::ndarray::f(x, y, (z), t): numpy arrays are read from
host memory (copied to the device), torch CUDA tensors are read in place.
Every call goes to the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

_contexts = {}


def context(device: int = 0):
    """Per-device ftk2h_ctx, created on first use."""
    ctx = _contexts.get(device)
    if ctx is None:
        p = C.c_void_p()
        _lib.check(_lib.lib().ftk2h_ctx_create(device, C.byref(p)))
        ctx = p
        _contexts[device] = ctx
    return ctx


class RegularSimplicialMesh:
    def __init__(self, dims, offset=None, global_dims=None):
        self.dims = [int(x) for x in dims]
        self.offset = [int(x) for x in offset] if offset else [0] * len(self.dims)
        self.global_dims = [int(x) for x in global_dims] if global_dims else list(self.dims)

    def get_total_dimension(self):
        return len(self.dims)

    def desc(self):
        d = _lib.MeshDesc()
        d.ndims = len(self.dims)
        for j in range(4):
            d.local_dims[j] = self.dims[j] if j < len(self.dims) else 1
            d.offset[j] = self.offset[j] if j < len(self.dims) else 0
            d.global_dims[j] = self.global_dims[j] if j < len(self.dims) else 1
        return d


class ExtrudedSimplicialMesh:
    """A RegularSimplicialMesh base stacked in n_layers + 1 time layers
    (mesh.hpp:542-560,711-726): the layers are one more grid axis and
    vertex id = spatial id + t * n_spatial_verts_, with n_spatial_verts_ =
    grid_index_to_id(base global dims) (larger than the spatial vertex count).
    Arrays are indexed [x, y, (z), t] at global coordinates, so the base must
    have no offset."""

    def __init__(self, base: RegularSimplicialMesh, n_layers: int):
        if any(base.offset):
            raise ValueError("ExtrudedSimplicialMesh: the base must have no offset")
        self.base = base
        self.n_layers = int(n_layers)
        self.dims = list(base.dims) + [self.n_layers + 1]
        self.offset = [0] * len(self.dims)
        self.global_dims = list(base.global_dims) + [self.n_layers + 1]
        r, mul = 0, 1
        for g in base.global_dims:
            r += g * mul
            mul *= g
        self.time_stride = r

    def get_total_dimension(self):
        return len(self.dims)

    def get_spatial_dimension(self):
        return len(self.base.dims)

    def desc(self):
        d = _lib.MeshDesc()
        d.ndims = len(self.dims)
        for j in range(4):
            d.local_dims[j] = self.dims[j] if j < len(self.dims) else 1
            d.offset[j] = 0
            d.global_dims[j] = self.global_dims[j] if j < len(self.dims) else 1
        d.time_id_stride = self.time_stride
        return d


@dataclass
class CriticalPointPredicate:
    M: int = 3
    var_names: list = field(default_factory=list)
    vector_var_name: str = ""
    scalar_var_name: str = ""
    use_multicomponent: bool = True
    sos_q: float = 1000000.0
    kind = _lib.PRED_CRITICAL_POINT

    @property
    def codimension(self):
        return self.M


@dataclass
class ContourPredicate:
    var_name: str = ""
    threshold: float = 0.0
    sos_q: float = 1000000.0
    kind = _lib.PRED_CONTOUR
    codimension = 1


@dataclass
class FiberPredicate:
    var_names: list = field(default_factory=lambda: ["", ""])
    thresholds: list = field(default_factory=lambda: [0.0, 0.0])
    sos_q: float = 1000000.0
    kind = _lib.PRED_FIBER
    codimension = 2


def _is_torch_cuda(a):
    try:
        import torch
        return isinstance(a, torch.Tensor) and a.is_cuda
    except ImportError:
        return False


def _dtype_of(a):
    if _is_torch_cuda(a):
        import torch
        if a.dtype == torch.float32:
            return _lib.F32
        if a.dtype == torch.float64:
            return _lib.F64
        raise TypeError(f"unsupported dtype {a.dtype}")
    if a.dtype == np.float32:
        return _lib.F32
    if a.dtype == np.float64:
        return _lib.F64
    raise TypeError(f"unsupported dtype {a.dtype}")


def field_view(a, shift=None):
    """ftk2h_field_view of an array indexed [x, y, (z), t] by local coordinates.
    shift: per-axis coordinate offset added to every index (multicomponent mode and
    scalar interpolation read global coordinates, engine.hpp:1108-1121)."""
    fv = _lib.FieldView()
    if _is_torch_cuda(a):
        st = list(a.stride())
        item = a.element_size()
        base = a.data_ptr()
        fv.on_device = 1
    else:
        st = [s // a.itemsize for s in a.strides]
        item = a.itemsize
        base = a.ctypes.data
        fv.on_device = 0
    if shift is not None:
        base += int(sum(int(o) * s for o, s in zip(shift, st))) * item
    fv.ptr = base
    for j in range(4):
        fv.strides[j] = st[j] if j < len(st) else 0
    return fv


class Result:
    """Owns an ftk2h_result handle."""

    def __init__(self, handle, keep=None):
        self.h = handle
        self._keep = keep

    def __del__(self):
        if getattr(self, "h", None):
            _lib.lib().ftk2h_result_free(self.h)
            self.h = None

    def sizes(self):
        n = C.c_uint64()
        m = C.c_int()
        nc = C.c_int()
        _lib.check(_lib.lib().ftk2h_result_sizes(self.h, C.byref(n), C.byref(m), C.byref(nc)))
        return n.value, m.value, nc.value

    def phase_ms(self):
        """Device ms per phase: k_classify, k_stage2, node scan, k_offsets, cell scan +
        sync, k_emit, k_link_pairs, k_compress (HIP events on the extraction's stream)."""
        buf = (C.c_float * 8)()
        _lib.check(_lib.lib().ftk2h_result_phase_ms(self.h, buf, 8))
        return list(buf)

    def device_view(self):
        dv = _lib.DeviceView()
        _lib.check(_lib.lib().ftk2h_result_device_view(self.h, C.byref(dv)))
        return dv

    def keys(self):
        n, m, _ = self.sizes()
        k = np.zeros(n, np.uint64)
        if n:
            _lib.check(_lib.lib().ftk2h_result_copy_keys(self.h, k.ctypes.data))
        return k

    def classify(self, mesh, comps, mode: int, device: int = 0):
        """FeatureElement::type of every critical point (ftk2h_result_classify): mode
        _lib.CLASSIFY_GRADIENT (Hessian eigenvalue signs) or _lib.CLASSIFY_FLOW
        (Jacobian eigenvalues); comps: the extraction's field components."""
        comps = list(comps)
        views = (_lib.FieldView * len(comps))(*[field_view(c) for c in comps])
        md = mesh.desc()
        _lib.check(_lib.lib().ftk2h_result_classify(context(device), self.h, C.byref(md), views, len(comps),
                                                    int(mode), None))
        return self.types()

    def types(self):
        n, m, _ = self.sizes()
        out = np.zeros(n, np.uint32)
        if n:
            _lib.check(_lib.lib().ftk2h_result_copy_types(self.h, out.ctypes.data))
        return out

    def add_attributes(self, mesh, specs, device: int = 0):
        """specs: [(slot, [arrays indexed by local coordinates])] (several arrays: norm)."""
        arr = (_lib.AttributeDesc * max(1, len(specs)))()
        for i, (slot, arrays) in enumerate(specs):
            arr[i].slot = slot
            arr[i].ncomp = len(arrays)
            for c, a in enumerate(arrays):
                arr[i].comps[c] = field_view(a)
        md = mesh.desc()
        _lib.check(_lib.lib().ftk2h_result_add_attributes(context(device), self.h, C.byref(md), arr, len(specs),
                                                          None))
        self._attr_keep = specs

    def attributes(self):
        n, _, _ = self.sizes()
        out = np.zeros((n, 16), np.float32)
        if n:
            _lib.check(_lib.lib().ftk2h_result_copy_attributes(self.h, out.ctypes.data))
        return out

    def node_range(self, begin: int, count: int):
        """(vertices, lambda, bary, scalar, track_id) of nodes [begin, begin+count)."""
        n, m, _ = self.sizes()
        nv = m + 1
        verts = np.zeros((count, nv), np.uint64)
        lam = np.zeros((count, nv), np.float64)
        bary = np.zeros((count, nv), np.float32)
        scal = np.zeros(count, np.float32)
        track = np.zeros(count, np.uint64)
        _lib.check(_lib.lib().ftk2h_result_copy_node_range(
            self.h, begin, count, verts.ctypes.data, lam.ctypes.data, bary.ctypes.data,
            scal.ctypes.data, track.ctypes.data))
        return verts, lam, bary, scal, track

    def positions(self, with_time=False):
        """(n, 3) float32 node positions as the reference's VTK writer computes them
        (computed on the device); with_time: also the "Time" array (the layer
        coordinate on an extruded mesh, else 0)."""
        n, _, _ = self.sizes()
        xyz = np.zeros((n, 3), np.float32)
        tm = np.zeros(n, np.float32)
        if n:
            _lib.check(_lib.lib().ftk2h_result_copy_positions(self.h, xyz.ctypes.data, tm.ctypes.data))
        return (xyz, tm) if with_time else xyz

    def track_ids(self):
        n, _, _ = self.sizes()
        track = np.zeros(n, np.uint64)
        if n:
            _lib.check(_lib.lib().ftk2h_result_copy_nodes(self.h, None, None, None, None,
                                                          track.ctypes.data))
        return track

    def lambdas(self):
        n, m, _ = self.sizes()
        lam = np.zeros((n, m + 1), np.float64)
        if n:
            _lib.check(_lib.lib().ftk2h_result_copy_nodes(self.h, None, lam.ctypes.data, None, None,
                                                          None))
        return lam

    def connectivity(self):
        _, _, nconn = self.sizes()
        conn = {}
        for i in range(nconn):
            dim = C.c_int()
            cells = C.c_uint64()
            _lib.check(_lib.lib().ftk2h_result_conn_info(self.h, i, C.byref(dim), C.byref(cells)))
            buf = np.zeros(cells.value * (dim.value + 1), np.uint32)
            if cells.value:
                _lib.check(_lib.lib().ftk2h_result_copy_conn(self.h, i, buf.ctypes.data))
            conn[dim.value] = buf.reshape(-1, dim.value + 1)
        return conn

    def to_complex(self):
        n, m, nconn = self.sizes()
        nv = m + 1
        verts = np.zeros((n, nv), np.uint64)
        lam = np.zeros((n, nv), np.float64)
        bary = np.zeros((n, nv), np.float32)
        scal = np.zeros(n, np.float32)
        track = np.zeros(n, np.uint64)
        if n:
            _lib.check(_lib.lib().ftk2h_result_copy_nodes(self.h, verts.ctypes.data, lam.ctypes.data,
                                                          bary.ctypes.data, scal.ctypes.data,
                                                          track.ctypes.data))
        return FeatureComplex(verts, lam, bary, scal, track, self.connectivity())


@dataclass
class FeatureComplex:
    vertices: np.ndarray      # (n, m+1) sorted vertex ids of each node's simplex
    lam: np.ndarray           # (n, m+1) solver output (double)
    bary: np.ndarray          # (n, m+1) FeatureElement::barycentric_coords[0] (float)
    scalar: np.ndarray        # (n,) FeatureElement::scalar
    track_id: np.ndarray      # (n,) minimum canonical index of the component
    connectivity: dict        # dim -> (cells, dim+1) uint32, sorted, unique


def extract(mesh: RegularSimplicialMesh, kind: int, m: int, comps, thresholds=(0.0, 0.0),
            sos_q: float = 1e6, scalar=None, scalar_shift=None, device: int = 0, stream=None,
            node_capacity: int = 0, cell_capacity: int = 0) -> Result:
    """One ftk2h_extract_regular call.  comps: arrays indexed by local coordinates."""
    comps = list(comps)
    dt = _dtype_of(comps[0])
    for c in comps:
        if _dtype_of(c) != dt:
            raise TypeError("all components must share one dtype")
    views = (_lib.FieldView * len(comps))(*[field_view(c) for c in comps])
    pd = _lib.PredicateDesc()
    pd.kind = kind
    pd.codimension = m
    pd.dtype = dt
    pd.sos_q = sos_q
    pd.thresholds[0] = float(thresholds[0])
    pd.thresholds[1] = float(thresholds[1]) if len(thresholds) > 1 else 0.0
    sv = None
    if scalar is not None:
        sv = field_view(scalar, scalar_shift)
    opt = _lib.Options(node_capacity, cell_capacity, 1)
    out = C.c_void_p()
    md = mesh.desc()
    _lib.check(_lib.lib().ftk2h_extract_regular(
        context(device), C.byref(md), C.byref(pd), views, len(comps),
        C.byref(sv) if sv is not None else None, C.byref(opt), C.byref(out),
        C.c_void_p(stream) if stream else None))
    return Result(out.value, keep=(comps, scalar))


class SimplicialEngine:
    """execute()/get_complex() on a RegularSimplicialMesh (engine.hpp:79-99, 822)."""

    def __init__(self, mesh: RegularSimplicialMesh, pred, device: int = 0):
        self.mesh = mesh
        self.pred = pred
        self.device = device
        self._result = None

    def execute(self, data: dict, var_names=None):
        p = self.pred
        mesh = self.mesh
        scalar = None
        shift = None
        if isinstance(p, CriticalPointPredicate):
            if p.use_multicomponent and p.vector_var_name:
                # [M, x, y, (z), t] read at GLOBAL coordinates (engine.hpp:1101-1124)
                vec = data[p.vector_var_name]
                comps = [vec[j] for j in range(p.M)]
                comps = [_shifted(c, mesh.offset) for c in comps]
            else:
                comps = [data[n] for n in p.var_names[: p.M]]
            if p.scalar_var_name:
                scalar = data[p.scalar_var_name]
                shift = mesh.offset  # predicate.hpp:84-89 reads global coordinates
            thr = (0.0, 0.0)
        elif isinstance(p, ContourPredicate):
            comps = [data[p.var_name]]
            thr = (p.threshold, 0.0)
        elif isinstance(p, FiberPredicate):
            comps = [data[p.var_names[0]], data[p.var_names[1]]]
            thr = tuple(p.thresholds)
        else:
            raise TypeError("unsupported predicate")
        self._result = extract(mesh, p.kind, p.codimension, comps, thr, p.sos_q, scalar, shift,
                               device=self.device)

    def get_complex(self) -> FeatureComplex:
        if self._result is None:
            raise RuntimeError("execute() first")
        return self._result.to_complex()


def _shifted(a, shift):
    if not any(shift):
        return a
    sl = tuple(slice(int(o), None) for o in shift)
    return a[sl]


class Stream:
    """Timestep-by-timestep extraction (ftk2h_stream_*): execute_stream semantics with
    the last timestep and the open track frontier resident on the GPU.  retire_bytes:
    device payload of finished nodes/lambda/cells beyond which they move to host
    memory (0: after every slab; None: the library default, 8 GiB).  window: timesteps
    per extraction (2: one slab per appended timestep, as the reference; W + 1: W
    slabs at once, the last partial window in finish())."""

    def __init__(self, spatial_dims, kind: int, m: int, dtype: int, ncomp: int, thresholds=(0.0, 0.0),
                 sos_q: float = 1e6, device: int = 0, retire_bytes=None, window: int = 2):
        self.spatial = [int(x) for x in spatial_dims]
        pd = _lib.PredicateDesc()
        pd.kind = kind
        pd.codimension = m
        pd.dtype = dtype
        pd.sos_q = sos_q
        pd.thresholds[0] = float(thresholds[0])
        pd.thresholds[1] = float(thresholds[1]) if len(thresholds) > 1 else 0.0
        sp = (C.c_uint64 * 3)(*(self.spatial + [1] * (3 - len(self.spatial))))
        h = C.c_void_p()
        _lib.check(_lib.lib().ftk2h_stream_open(context(device), len(self.spatial), sp, C.byref(pd), ncomp,
                                                C.byref(h)))
        self.h = h.value
        self.ncomp = ncomp
        if retire_bytes is not None:
            _lib.check(_lib.lib().ftk2h_stream_set_retire_bytes(self.h, int(retire_bytes)))
        if window != 2:
            _lib.check(_lib.lib().ftk2h_stream_set_window(self.h, int(window)))

    def resident_bytes(self) -> int:
        """Device bytes in use (timestep buffers + committed state still on the GPU)."""
        return int(_lib.lib().ftk2h_stream_resident_bytes(self.h))

    def retired_nodes(self) -> int:
        return int(_lib.lib().ftk2h_stream_retired_nodes(self.h))

    def append(self, comps):
        comps = list(comps)
        if len(comps) != self.ncomp:
            raise ValueError("wrong number of components")
        views = (_lib.FieldView * len(comps))(*[field_view(c) for c in comps])
        _lib.check(_lib.lib().ftk2h_stream_append(self.h, views, None))

    def finish(self) -> Result:
        out = C.c_void_p()
        _lib.check(_lib.lib().ftk2h_stream_finish(self.h, C.byref(out)))
        return Result(out.value)

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ftk2h_stream_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


def _pred_desc(kind, m, dtype, thresholds=(0.0, 0.0), sos_q=1e6):
    pd = _lib.PredicateDesc()
    pd.kind = kind
    pd.codimension = m
    pd.dtype = dtype
    pd.sos_q = sos_q
    pd.thresholds[0] = float(thresholds[0])
    pd.thresholds[1] = float(thresholds[1]) if len(thresholds) > 1 else 0.0
    return pd


class Batch:
    """Many meshes of one predicate in one pass (ftk2h_batch_*).  run() takes a list of
    (mesh, comps) or (mesh, comps, scalar) and returns one Result per job, each equal
    to extract() on that job alone."""

    def __init__(self, kind: int, m: int, dtype: int, ndims: int, thresholds=(0.0, 0.0), sos_q: float = 1e6,
                 device: int = 0):
        self.pd = _pred_desc(kind, m, dtype, thresholds, sos_q)
        h = C.c_void_p()
        _lib.check(_lib.lib().ftk2h_batch_create(context(device), C.byref(self.pd), ndims, C.byref(h)))
        self.h = h.value

    def _jobs(self, jobs):
        arr = (_lib.Job * len(jobs))()
        keep = []
        for i, job in enumerate(jobs):
            mesh, comps = job[0], list(job[1])
            scalar = job[2] if len(job) > 2 else None
            views = (_lib.FieldView * len(comps))(*[field_view(c) for c in comps])
            keep.append((views, comps, scalar))
            arr[i].mesh = mesh.desc()
            arr[i].comps = C.cast(views, C.POINTER(_lib.FieldView))
            if scalar is not None:
                sv = _lib.FieldView * 1
                sva = sv(field_view(scalar))
                keep.append(sva)
                arr[i].scalar = C.cast(sva, C.POINTER(_lib.FieldView))
        return arr, keep

    def set_config(self, grid_per_cu=0, chunk=0, use_graph=True):
        cfg = _lib.BatchConfig(grid_per_cu, chunk, 1 if use_graph else 0)
        _lib.check(_lib.lib().ftk2h_batch_set_config(self.h, C.byref(cfg)))

    def config(self):
        cfg = _lib.BatchConfig()
        _lib.check(_lib.lib().ftk2h_batch_get_config(self.h, C.byref(cfg)))
        return {"grid_per_cu": cfg.grid_per_cu, "chunk": cfg.chunk, "use_graph": cfg.use_graph}

    def run(self, jobs, stream=None):
        arr, keep = self._jobs(jobs)
        outs = (C.c_void_p * len(jobs))()
        _lib.check(_lib.lib().ftk2h_batch_run(self.h, len(jobs), arr, outs,
                                              C.c_void_p(stream) if stream else None))
        return [Result(outs[i], keep=keep) for i in range(len(jobs))]

    def autotune(self, jobs, reps=2):
        arr, keep = self._jobs(jobs)
        cfg = _lib.BatchConfig()
        ms = C.c_float()
        _lib.check(_lib.lib().ftk2h_batch_autotune(self.h, len(jobs), arr, reps, C.byref(cfg), C.byref(ms)))
        return {"grid_per_cu": cfg.grid_per_cu, "chunk": cfg.chunk, "use_graph": cfg.use_graph}, ms.value

    def last_ms(self):
        buf = (C.c_float * 3)()
        _lib.check(_lib.lib().ftk2h_batch_last_ms(self.h, buf, 3))
        return list(buf)

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ftk2h_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Server:
    """Continuous-batching request server (ftk2h_server_*): streams admitted against a
    device-memory budget; step() advances every admitted stream by one slab, the
    slabs of one predicate template in one batched pass."""

    def __init__(self, budget_bytes: int, device: int = 0):
        h = C.c_void_p()
        _lib.check(_lib.lib().ftk2h_server_create(context(device), int(budget_bytes), C.byref(h)))
        self.h = h.value
        self._keep = {}

    def open(self, spatial_dims, kind: int, m: int, dtype: int, ncomp: int, max_timesteps: int,
             thresholds=(0.0, 0.0), sos_q: float = 1e6):
        """Stream id, or None when the budget is in use (retry after taking results)."""
        pd = _pred_desc(kind, m, dtype, thresholds, sos_q)
        sp = [int(x) for x in spatial_dims]
        dims = (C.c_uint64 * 3)(*(sp + [1] * (3 - len(sp))))
        sid = C.c_uint64()
        rc = _lib.lib().ftk2h_server_open(self.h, len(sp), dims, C.byref(pd), ncomp, int(max_timesteps),
                                          C.byref(sid))
        if rc == _lib.E_AGAIN:
            return None
        _lib.check(rc)
        self._keep[sid.value] = []
        return sid.value

    def append(self, sid: int, comps):
        comps = list(comps)
        views = (_lib.FieldView * len(comps))(*[field_view(c) for c in comps])
        _lib.check(_lib.lib().ftk2h_server_append(self.h, sid, views))
        self._keep[sid].append(comps)  # device views are read by a later step

    def close_stream(self, sid: int):
        _lib.check(_lib.lib().ftk2h_server_close(self.h, sid))

    def step(self, max_done: int = 4096):
        done = (C.c_uint64 * max_done)()
        n = C.c_int()
        _lib.check(_lib.lib().ftk2h_server_step(self.h, done, max_done, C.byref(n)))
        return [done[i] for i in range(min(n.value, max_done))]

    def take(self, sid: int) -> Result:
        out = C.c_void_p()
        _lib.check(_lib.lib().ftk2h_server_take(self.h, sid, C.byref(out)))
        self._keep.pop(sid, None)
        return Result(out.value)

    def stats(self):
        r = C.c_uint64()
        d = C.c_uint64()
        a = C.c_int()
        q = C.c_int()
        _lib.check(_lib.lib().ftk2h_server_stats(self.h, C.byref(r), C.byref(d), C.byref(a), C.byref(q)))
        return {"reserved_bytes": r.value, "device_bytes": d.value, "active": a.value, "queued": q.value}

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().ftk2h_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
