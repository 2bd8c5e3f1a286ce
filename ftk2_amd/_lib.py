"""ctypes bindings of libftk2h.so (include/ftk2h.h).

The library is loaded from the package directory only; there is no fallback: if
it is missing, or no GPU is visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libftk2h.so")

FTK2H_OK = 0
STATUS = {0: "OK", 1: "E_INVALID", 2: "E_UNSUPPORTED", 3: "E_OVERFLOW", 4: "E_OOM", 5: "E_HIP", 6: "E_AGAIN"}
E_AGAIN = 6
PRED_CRITICAL_POINT, PRED_CONTOUR, PRED_FIBER = 0, 1, 2
CLASSIFY_GRADIENT, CLASSIFY_FLOW = 1, 2
CP_TYPES = {0: "unclassified", 1: "minimum", 2: "maximum", 3: "saddle", 4: "saddle-1", 5: "saddle-2",
            16: "source", 17: "sink", 18: "saddle", 19: "spiral source", 20: "spiral sink", 21: "center",
            255: "degenerate"}
F32, F64 = 0, 1


class Ftk2hError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("ndims", C.c_int), ("local_dims", C.c_uint64 * 4), ("offset", C.c_uint64 * 4),
                ("global_dims", C.c_uint64 * 4), ("time_id_stride", C.c_uint64)]


class PredicateDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("codimension", C.c_int), ("dtype", C.c_int), ("sos_q", C.c_double),
                ("thresholds", C.c_double * 2)]


class FieldView(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("strides", C.c_int64 * 4), ("on_device", C.c_int)]


class Options(C.Structure):
    _fields_ = [("node_capacity", C.c_uint64), ("cell_capacity", C.c_uint64), ("sync", C.c_int)]


class DeviceView(C.Structure):
    _fields_ = [("n_nodes", C.c_uint64), ("node_dim", C.c_int), ("keys", C.c_void_p),
                ("lambda_", C.c_void_p), ("track_id", C.c_void_p), ("scalar", C.c_void_p),
                ("n_conn", C.c_int), ("conn_dim", C.c_int * 3), ("conn_cells", C.c_uint64 * 3),
                ("conn", C.c_void_p * 3), ("scalar_value", C.c_float), ("n_types", C.c_int),
                ("global_dims", C.c_uint64 * 4), ("time_id_stride", C.c_uint64)]


class AttributeDesc(C.Structure):
    _fields_ = [("slot", C.c_int), ("ncomp", C.c_int), ("comps", FieldView * 16)]


class Job(C.Structure):
    _fields_ = [("mesh", MeshDesc), ("comps", C.POINTER(FieldView)), ("scalar", C.POINTER(FieldView))]


class OutputBuffers(C.Structure):
    _fields_ = [("keys", C.c_void_p), ("lambda_", C.c_void_p), ("track_id", C.c_void_p), ("scalar", C.c_void_p),
                ("cells", C.c_void_p), ("node_capacity", C.c_uint64), ("cell_capacity", C.c_uint64)]


class BatchConfig(C.Structure):
    _fields_ = [("grid_per_cu", C.c_int), ("chunk", C.c_int), ("use_graph", C.c_int)]


# Every entry point of include/ftk2h.h: name -> (restype, argtypes)
_P = C.c_void_p
SIGNATURES = {
    "ftk2h_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "ftk2h_ctx_destroy": (None, [_P]),
    "ftk2h_extract_regular": (C.c_int, [_P, C.POINTER(MeshDesc), C.POINTER(PredicateDesc),
                                        C.POINTER(FieldView), C.c_int, C.POINTER(FieldView),
                                        C.POINTER(Options), C.POINTER(_P), _P]),
    "ftk2h_result_sizes": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ftk2h_result_conn_info": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_uint64)]),
    "ftk2h_result_copy_nodes": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "ftk2h_result_copy_positions": (C.c_int, [_P, _P, _P]),
    "ftk2h_result_copy_node_range": (C.c_int, [_P, C.c_uint64, C.c_uint64, _P, _P, _P, _P, _P]),
    "ftk2h_result_copy_keys": (C.c_int, [_P, _P]),
    "ftk2h_result_copy_conn": (C.c_int, [_P, C.c_int, _P]),
    "ftk2h_result_device_view": (C.c_int, [_P, C.POINTER(DeviceView)]),
    "ftk2h_result_phase_ms": (C.c_int, [_P, C.POINTER(C.c_float), C.c_int]),
    "ftk2h_result_free": (None, [_P]),
    "ftk2h_last_error": (C.c_char_p, []),
    "ftk2h_stream_open": (C.c_int, [_P, C.c_int, C.POINTER(C.c_uint64), C.POINTER(PredicateDesc), C.c_int,
                                    C.POINTER(_P)]),
    "ftk2h_stream_append": (C.c_int, [_P, C.POINTER(FieldView), _P]),
    "ftk2h_stream_finish": (C.c_int, [_P, C.POINTER(_P)]),
    "ftk2h_stream_close": (None, [_P]),
    "ftk2h_stream_set_retire_bytes": (C.c_int, [_P, C.c_uint64]),
    "ftk2h_stream_set_window": (C.c_int, [_P, C.c_int]),
    "ftk2h_stream_resident_bytes": (C.c_uint64, [_P]),
    "ftk2h_stream_retired_nodes": (C.c_uint64, [_P]),
    "ftk2h_batch_create": (C.c_int, [_P, C.POINTER(PredicateDesc), C.c_int, C.POINTER(_P)]),
    "ftk2h_batch_set_config": (C.c_int, [_P, C.POINTER(BatchConfig)]),
    "ftk2h_batch_get_config": (C.c_int, [_P, C.POINTER(BatchConfig)]),
    "ftk2h_batch_run": (C.c_int, [_P, C.c_int, C.POINTER(Job), C.POINTER(_P), _P]),
    "ftk2h_batch_autotune": (C.c_int, [_P, C.c_int, C.POINTER(Job), C.c_int, C.POINTER(BatchConfig),
                                       C.POINTER(C.c_float)]),
    "ftk2h_batch_last_ms": (C.c_int, [_P, C.POINTER(C.c_float), C.c_int]),
    "ftk2h_batch_destroy": (None, [_P]),
    "ftk2h_result_add_attributes": (C.c_int, [_P, _P, C.POINTER(MeshDesc), C.POINTER(AttributeDesc), C.c_int,
                                              _P]),
    "ftk2h_result_copy_attributes": (C.c_int, [_P, _P]),
    "ftk2h_result_classify": (C.c_int, [_P, _P, C.POINTER(MeshDesc), C.POINTER(FieldView), C.c_int, C.c_int, _P]),
    "ftk2h_result_copy_types": (C.c_int, [_P, _P]),
    "ftk2h_server_create": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "ftk2h_server_open": (C.c_int, [_P, C.c_int, C.POINTER(C.c_uint64), C.POINTER(PredicateDesc), C.c_int,
                                    C.c_uint64, C.POINTER(C.c_uint64)]),
    "ftk2h_server_append": (C.c_int, [_P, C.c_uint64, C.POINTER(FieldView)]),
    "ftk2h_server_close": (C.c_int, [_P, C.c_uint64]),
    "ftk2h_server_step": (C.c_int, [_P, C.POINTER(C.c_uint64), C.c_int, C.POINTER(C.c_int)]),
    "ftk2h_server_take": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "ftk2h_server_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "ftk2h_server_destroy": (None, [_P]),
    "ftk2h_extract_regular_noalloc": (C.c_int, [_P, C.POINTER(MeshDesc), C.POINTER(PredicateDesc),
                                                C.POINTER(FieldView), C.c_int, C.POINTER(FieldView),
                                                C.POINTER(OutputBuffers), C.POINTER(C.c_uint64),
                                                C.POINTER(C.c_uint64), _P]),
}

_lib = None


def lib():
    """Load libftk2h.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ftk2_amd.build` "
                               "(there is no CPU fallback)")
        # One HIP runtime per process.  PyTorch-ROCm carries its own libamdhip64 and
        # libhsa-runtime64 (SONAMEs libamdhip64.so.7 / libhsa-runtime64.so.1, but its
        # libraries ask for them as libamdhip64.so / libhsa-runtime64.so).  Loaded first,
        # this library would bind the system copies, a later `import torch` would map a
        # second HIP and HSA runtime beside them, and torch's device initialisation would
        # then find no GPU.  With torch imported first, libamdhip64.so.7 resolves by its
        # SONAME to the runtime torch already mapped.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != FTK2H_OK:
        raise Ftk2hError(code, lib().ftk2h_last_error().decode(errors="replace"))
