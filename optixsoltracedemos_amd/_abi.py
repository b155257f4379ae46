"""ctypes mirror of include/soltrace_hip.h.

Loads the in-tree HIP library ``optixsoltracedemos_amd/_lib/libsoltrace_hip.so``
(built by ``__graft_entry__.build()``). There is no fallback: if the library is
missing, :func:`load` raises, so no code path silently runs on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "_lib")
# ST_HIP_LIB selects another build of the library (A/B timing of two builds)
LIB_PATH = os.environ.get("ST_HIP_LIB") or os.path.join(LIB_DIR, "libsoltrace_hip.so")

ST_OK = 0
ST_ERR_INVALID = -1
ST_ERR_DEVICE = -2
ST_ERR_ALLOC = -3
ST_ERR_STATE = -4
ST_ERR_REJECTED = -5
ST_ERR_CANCELLED = -6
ST_ERR_INJECTED = -7

SURFACE_FLAT, SURFACE_PARABOLIC, SURFACE_CYLINDER = 0, 1, 3
APERTURE_RECTANGLE, APERTURE_CIRCLE = 0, 1
ROLE_MIRROR, ROLE_RECEIVER = 0, 1
SUN_PILLBOX, SUN_GAUSSIAN = 0, 1
POS_PHILOX, POS_HALTON, POS_STRATIFIED = 0, 1, 2

TRACE_RECORDS = 0x1
TRACE_FLUX_DIRECT_SUN = 0x2
TRACE_ZERO_OUTPUTS = 0x4
TRACE_BVH_ONLY = 0x8
TRACE_COUNT_WORK = 0x10

KERNEL_AUTO, KERNEL_MEGA, KERNEL_QUEUED, KERNEL_QUEUED_GROUPED = 0, 1, 2, 3

STAT_NAMES = [
    "rays", "segments", "mirror_hits", "receiver_hits", "direct_sun", "absorbed",
    "escaped", "back_face", "truncated", "cap_hits", "flux_binned",
    "node_visits", "flat_tests", "parabolic_tests", "cylinder_tests",
]
STAT_COUNT = 16


class Element(C.Structure):
    """st_element_f64 (FP64 element: lib/element.h:41-102 + Surface + Aperture)."""
    _fields_ = [
        ("origin", C.c_double * 3),
        ("aim_point", C.c_double * 3),
        ("zrot_deg", C.c_double),
        ("dim_x", C.c_double),
        ("dim_y", C.c_double),
        ("curv_x", C.c_double),
        ("curv_y", C.c_double),
        ("reflectivity", C.c_double),
        ("slope_error", C.c_double),
        ("spec_error", C.c_double),
        ("surface", C.c_int32),
        ("aperture", C.c_int32),
        ("role", C.c_int32),
        ("reserved", C.c_int32),
    ]


ELEMENT_DTYPE = np.dtype([
    ("origin", "<f8", (3,)), ("aim_point", "<f8", (3,)), ("zrot_deg", "<f8"),
    ("dim_x", "<f8"), ("dim_y", "<f8"), ("curv_x", "<f8"), ("curv_y", "<f8"),
    ("reflectivity", "<f8"), ("slope_error", "<f8"), ("spec_error", "<f8"),
    ("surface", "<i4"), ("aperture", "<i4"), ("role", "<i4"), ("reserved", "<i4"),
])
assert ELEMENT_DTYPE.itemsize == C.sizeof(Element) == 128


class SunDesc(C.Structure):
    _fields_ = [
        ("vector", C.c_double * 3),
        ("param", C.c_double),
        ("dni", C.c_double),
        ("shape", C.c_int32),
        ("position_sampling", C.c_int32),
    ]


class SunPlane(C.Structure):
    _fields_ = [
        ("v0", C.c_double * 3), ("v1", C.c_double * 3), ("v2", C.c_double * 3), ("v3", C.c_double * 3),
        ("distance", C.c_double), ("area", C.c_double), ("box_half_angle", C.c_double),
    ]


class TraceDesc(C.Structure):
    _fields_ = [
        ("num_rays", C.c_uint64),
        ("ray_begin", C.c_uint64),
        ("ray_count", C.c_uint64),
        ("seed", C.c_uint64),
        ("job_id", C.c_uint32),
        ("max_depth", C.c_int32),
        ("flags", C.c_uint32),
        ("kernel", C.c_int32),
        ("flux_element", C.c_int32),
        ("flux_nx", C.c_int32),
        ("flux_ny", C.c_int32),
        ("block_size", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("waves_per_simd", C.c_int32),
        ("stream", C.c_void_p),
    ]


class Outputs(C.Structure):
    _fields_ = [
        ("flux", C.c_void_p),
        ("element_hits", C.c_void_p),
        ("stats", C.c_void_p),
        ("records", C.c_void_p),
        ("record_elements", C.c_void_p),
    ]


class BatchDesc(C.Structure):
    _fields_ = [
        ("steps", C.c_int32),
        ("max_depth", C.c_int32),
        ("sun_vectors", C.c_void_p),
        ("budgets", C.c_void_p),
        ("target", C.c_double * 3),
        ("sun", SunDesc),
        ("seed", C.c_uint64),
        ("flux_element", C.c_int32),
        ("flux_nx", C.c_int32),
        ("flux_ny", C.c_int32),
        ("flags", C.c_uint32),
        ("use_graph", C.c_int32),
        ("block_size", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("stream", C.c_void_p),
    ]


class TuneResult(C.Structure):
    _fields_ = [("block", C.c_int32), ("grid", C.c_int32), ("stack", C.c_int32), ("waves", C.c_int32),
                ("trials", C.c_int32), ("queued", C.c_int32), ("best_ms", C.c_double), ("default_ms", C.c_double)]


class LaunchInfo(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("block_size", C.c_int32), ("grid_blocks", C.c_int32), ("stack", C.c_int32),
                ("waves_per_simd", C.c_int32), ("chunk_rays", C.c_int32), ("lds_hits", C.c_int32),
                ("flux_copies", C.c_int32), ("chunks", C.c_uint64)]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("max_active_jobs", C.c_int32),
        ("chunk_rays", C.c_uint64),
        ("block_size", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("round_rays", C.c_uint64),
    ]


class JobDesc(C.Structure):
    _fields_ = [
        ("scene_slot", C.c_int32),
        ("sun", SunDesc),
        ("num_rays", C.c_uint64),
        ("seed", C.c_uint64),
        ("job_id", C.c_uint32),
        ("max_depth", C.c_int32),
        ("flux_nx", C.c_int32),
        ("flux_ny", C.c_int32),
        ("flags", C.c_uint32),
    ]


class JobResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("done", C.c_int32),
        ("submit_s", C.c_double),
        ("start_s", C.c_double),
        ("finish_s", C.c_double),
        ("power_per_ray", C.c_double),
        ("stats", C.c_uint64 * STAT_COUNT),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("admitted", C.c_uint64), ("rejected", C.c_uint64),
        ("completed", C.c_uint64), ("cancelled", C.c_uint64),
        ("rays_traced", C.c_uint64), ("launches", C.c_uint64),
        ("resident_bytes", C.c_uint64), ("peak_resident_bytes", C.c_uint64),
        ("busy_s", C.c_double),
    ]


# name -> (restype, argtypes); every symbol include/soltrace_hip.h declares
P = C.c_void_p
SIGNATURES = {
    "st_abi_version": (C.c_int, []),
    "st_last_error": (C.c_char_p, []),
    "st_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "st_device_memory": (C.c_int, [C.c_int, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "st_context_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "st_context_destroy": (C.c_int, [P]),
    "st_buffer_alloc": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
    "st_buffer_free": (C.c_int, [P, P]),
    "st_buffer_upload": (C.c_int, [P, P, P, C.c_size_t]),
    "st_buffer_download": (C.c_int, [P, P, P, C.c_size_t]),
    "st_buffer_zero": (C.c_int, [P, P, C.c_size_t]),
    "st_scene_create": (C.c_int, [P, C.POINTER(Element), C.c_int32, C.POINTER(P)]),
    "st_scene_destroy": (C.c_int, [P]),
    "st_scene_num_elements": (C.c_int, [P, C.POINTER(C.c_int32)]),
    "st_scene_update_aim": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "st_scene_update_elements": (C.c_int, [P, C.POINTER(Element), C.c_int32]),
    "st_scene_update": (C.c_int, [P, C.POINTER(Element), C.c_int32, C.POINTER(SunDesc)]),
    "st_scene_aim_heliostats": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "st_scene_set_sun": (C.c_int, [P, C.POINTER(SunDesc)]),
    "st_scene_set_sun_and_aim": (C.c_int, [P, C.POINTER(SunDesc), C.POINTER(C.c_double)]),
    "st_scene_get_sun_plane": (C.c_int, [P, C.POINTER(SunPlane)]),
    "st_scene_get_frames": (C.c_int, [P, C.POINTER(C.c_double)]),
    "st_scene_get_aabbs": (C.c_int, [P, C.POINTER(C.c_float)]),
    "st_scene_bvh_info": (C.c_int, [P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "st_scene_synchronize": (C.c_int, [P]),
    "st_scene_autotune": (C.c_int, [P, C.c_uint64, C.c_int32, C.POINTER(TuneResult)]),
    "st_scene_build_lists": (C.c_int, [P]),
    "st_scene_list_info": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "st_trace": (C.c_int, [P, C.POINTER(TraceDesc), C.POINTER(Outputs)]),
    "st_scene_last_launch": (C.c_int, [P, C.POINTER(LaunchInfo)]),
    "st_trace_noalloc": (C.c_int, [P, C.POINTER(TraceDesc), C.POINTER(Outputs)]),
    "st_flux_to_watts": (C.c_int, [P, C.POINTER(TraceDesc), P, P, C.POINTER(C.c_double)]),
    "st_trace_batch": (C.c_int, [P, C.POINTER(BatchDesc), P, P, P]),
    "st_server_create": (C.c_int, [P, C.POINTER(ServerConfig), C.POINTER(P)]),
    "st_server_destroy": (C.c_int, [P]),
    "st_server_add_scene": (C.c_int, [P, P, C.POINTER(C.c_int32)]),
    "st_server_submit": (C.c_int, [P, C.POINTER(JobDesc), C.POINTER(C.c_int64)]),
    "st_server_poll": (C.c_int, [P, C.c_int64, C.POINTER(JobResult)]),
    "st_server_wait": (C.c_int, [P, C.c_int64, C.POINTER(JobResult)]),
    "st_server_fetch": (C.c_int, [P, C.c_int64, P, P]),
    "st_server_cancel": (C.c_int, [P, C.c_int64]),
    "st_server_release": (C.c_int, [P, C.c_int64]),
    "st_server_get_stats": (C.c_int, [P, C.POINTER(ServerStats)]),
    "st_calibrate_valu": (C.c_int, [P, C.POINTER(C.c_double)]),
    "st_debug_inject_fault": (C.c_int, [C.c_int32, C.c_int32]),
}

_lib = None


class STError(RuntimeError):
    def __init__(self, name, code, msg):
        super().__init__(f"{name} failed ({code}): {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load the HIP library and declare every C-ABI signature. Raises if absent."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build the HIP extension first (python -c 'import __graft_entry__ as g; g.build()')")
    # torch bundles its own HIP runtime under the same soname (libamdhip64.so.7)
    # as the ROCm one this library links; the first one loaded serves the whole
    # process. Loading torch first keeps torch on the runtime it was built for
    # (the other order leaves torch reporting no GPU).
    import torch  # noqa: F401
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def check(name: str, code: int):
    if code != ST_OK:
        msg = _lib.st_last_error().decode() if _lib is not None else ""
        raise STError(name, code, msg)


def call(name: str, *args):
    lib = load()
    code = getattr(lib, name)(*args)
    check(name, code)
    return code


def elements_to_ctypes(elems: np.ndarray):
    """numpy ELEMENT_DTYPE array -> (ctypes array, n) sharing the bytes."""
    arr = np.ascontiguousarray(elems, dtype=ELEMENT_DTYPE)
    n = len(arr)
    carr = (Element * n).from_buffer_copy(arr.tobytes())
    return carr, n


def sun_desc(vector, param, shape=SUN_PILLBOX, dni=1000.0, position_sampling=POS_PHILOX) -> SunDesc:
    s = SunDesc()
    for k in range(3):
        s.vector[k] = float(vector[k])
    s.param = float(param)
    s.dni = float(dni)
    s.shape = int(shape)
    s.position_sampling = int(position_sampling)
    return s
