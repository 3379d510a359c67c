"""ctypes binding of libccd_amd.so (the C ABI declared in include/ccd_amd.h).

The library is the only compute path: if it is missing or cannot be loaded the import of the
functions below raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libccd_amd.so")

CCD_OK = 0
CCD_E_INVALID = 1
CCD_E_CAPACITY = 2
CCD_E_OOM = 3
CCD_E_HIP = 4
CCD_E_BUDGET = 5
CCD_E_NODEVICE = 6
NOALLOC_MAX_STENCILS = 1 << 20   # CCD_NOALLOC_MAX_STENCILS

PRIM_VERTEX_FACE = 0
PRIM_EDGE_EDGE = 1
PRIM_VERTEX_EDGE = 2
PRIM_VERTEX_VERTEX = 3

CCD_SORTED = 1
CCD_GRAPH = 2

DIST_VERTEX_PLANE_LT = 0
DIST_LINE_LINE_LT = 1
DIST_VERTEX_FACE = 2
DIST_EDGE_EDGE = 3

# Every exported symbol of include/ccd_amd.h with its ctypes signature.
_V = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_INT = ctypes.c_int
_P64 = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = {
    "ccd_create": ([ctypes.POINTER(_V), _INT], _INT),
    "ccd_destroy": ([_V], _INT),
    "ccd_set_stream": ([_V, _V], _INT),
    "ccd_synchronize": ([_V], _INT),
    "ccd_last_error": ([_V], ctypes.c_char_p),
    "ccd_device_bytes": ([_V], _I64),
    "ccd_ctcd_primitives": ([_V, _INT, _V, _V, _D, _I64, _V, _V], _INT),
    "ccd_ctcd_narrowphase": ([_V, _V, _V, _I64, _V, _V, _I64, _V, _V, _I64, _D, _V, _V, _V, _V, _V, _V], _INT),
    "ccd_ctcd_narrowphase_noalloc": ([_V, _V, _V, _I64, _V, _V, _I64, _V, _V, _I64, _D, _V, _V, _V, _V, _V, _V],
                                     _INT),
    "ccd_separating_plane_narrowphase": ([_V, _V, _V, _V, _I64, _I64, _V, _V, _I64, _V, _V, _I64, _D, _V, _V, _P64],
                                         _INT),
    "ccd_ctcd_narrowphase_history": ([_V, _V, _V, _V, _I64, _I64, _V, _V, _I64, _V, _V, _I64, _D, _V, _V, _V, _V,
                                      _V, _V, _V, _V], _INT),
    "ccd_kdop_candidates": ([_V, _V, _I64, _V, _V, _I64, _D, _V, _V, _I64, _P64, _V, _I64, _P64, _INT], _INT),
    "ccd_kdop_candidates_history": ([_V, _V, _I64, _V, _V, _I64, _I64, _D, _V, _V, _I64, _P64, _V, _I64, _P64,
                                     _INT], _INT),
    "ccd_detect": ([_V, _V, _I64, _V, _V, _I64, _D, _D, _V, _V, _V, _I64, _P64, _V, _V, _I64, _P64, _P64, _P64],
                   _INT),
    "ccd_distance_primitives": ([_V, _INT, _V, _D, _I64, _V, _V, _V], _INT),
    "ccd_mesh_self_distance": ([_V, _V, _I64, _V, _I64, _V, ctypes.POINTER(_D), _P64, _P64], _INT),
    "ccd_device_alloc": ([_V, _I64, ctypes.POINTER(_V)], _INT),
    "ccd_device_free": ([_V, _V], _INT),
    "ccd_copy_to_device": ([_V, _V, _V, _I64], _INT),
    "ccd_copy_to_host": ([_V, _V, _V, _I64], _INT),
    "ccd_batch_detect": ([_V, _I64, _V, _V, _I64, _V, _V, _I64, _D, _D, _V, _V, _I64, _P64, _V, _V, _I64, _P64, _P64,
                          _P64, _INT], _INT),
    "ccd_autotune": ([_V, _V, _V, _D, ctypes.POINTER(_INT)], _INT),
    "ccd_set_queue_shape": ([_V, ctypes.POINTER(_INT)], _INT),
    "ccd_set_profiling": ([_V, _INT], _INT),
    "ccd_set_edge_edge_split": ([_V, _INT], _INT),
    "ccd_stage_times": ([_V, ctypes.POINTER(_D), _P64, _INT], _INT),
    "ccd_narrowphase_work": ([_V, _V, _V, _I64, _V, _I64, _V, _I64, _D, _P64, _P64], _INT),
    "ccd_stage_flops": ([_P64], _INT),
    "ccd_last_candidates": ([_V, ctypes.POINTER(_V), _P64, ctypes.POINTER(_V), _P64], _INT),
}

NTIMERS = 9
TIMER_NAMES = ("bounds", "sort", "build", "tables", "traverse", "expand", "narrow_vf", "narrow_ee", "pipeline")
NSTAGES = 12
STAGE_NAMES = ("vf_call", "vf_plane", "vf_dist", "ee_dist", "ee_par", "ee_bary", "ve_call", "ve_quad", "ve_quart",
               "vv", "quad_root", "isolate")

class RequestResult(ctypes.Structure):
    """ccd_request_result (include/ccd_amd.h)."""
    _fields_ = [("request_id", ctypes.c_int64), ("session", ctypes.c_int32), ("status", ctypes.c_int32),
                ("vf_hits", ctypes.c_int64), ("ee_hits", ctypes.c_int64), ("vf_candidates", ctypes.c_int64),
                ("ee_candidates", ctypes.c_int64), ("checksum_vf", ctypes.c_uint64), ("checksum_ee", ctypes.c_uint64),
                ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("reruns", ctypes.c_int32), ("reserved", ctypes.c_int32), ("face_pairs", ctypes.c_int64)]


class VFStats(ctypes.Structure):
    """ccd_vf_stats (include/ccd_amd.h)."""
    _fields_ = [("iterations", ctypes.c_int64), ("layers", ctypes.c_int64), ("firings", ctypes.c_int64),
                ("history_entries", ctypes.c_int64), ("vf_candidates", ctypes.c_int64),
                ("ee_candidates", ctypes.c_int64), ("vf_collisions", ctypes.c_int64),
                ("ee_collisions", ctypes.c_int64), ("closest", ctypes.c_double), ("host_seconds", ctypes.c_double),
                ("device_seconds", ctypes.c_double)]


SIGNATURES.update({
    "ccd_velocity_filter": ([_V, _V, _V, _I64, _V, _I64, _V, _D, _D, _D, _D, _D, _INT, _INT, ctypes.POINTER(_INT),
                             ctypes.POINTER(VFStats)], _INT),
    "ccd_server_create": ([_INT, _I64, _INT, ctypes.POINTER(_V)], _INT),
    "ccd_server_open_session": ([_V, _V, _I64, _V, _V, _I64, _D, _D, ctypes.POINTER(_INT)], _INT),
    "ccd_server_submit": ([_V, _INT, _V, _I64], _INT),
    "ccd_server_poll": ([_V, ctypes.POINTER(RequestResult), _INT, ctypes.POINTER(_INT)], _INT),
    "ccd_server_drain": ([_V], _INT),
    "ccd_server_stats": ([_V, _P64], _INT),
    "ccd_server_keep_hits": ([_V, _INT], _INT),
    "ccd_server_fetch_hits": ([_V, _I64, _V, _I64, _P64], _INT),
    "ccd_server_destroy": ([_V], _INT),
})

_lib = None


class CCDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ccd error {code}: {msg}")
        self.code = code


def lib():
    """Load libccd_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with collisiondetection_amd._build.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib
