"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ST_ORACLE_LIB selects another build of the same source (the sanitizer test's)
LIB = os.environ.get("ST_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

STATS = 16


class OrHit(C.Structure):
    _fields_ = [("t", C.c_double), ("p", C.c_double * 3), ("n", C.c_double * 3), ("local", C.c_double * 3),
                ("elem", C.c_int32), ("cap", C.c_int32)]


class OrSun(C.Structure):
    _fields_ = [("vector", C.c_double * 3), ("param", C.c_double), ("dni", C.c_double),
                ("shape", C.c_int32), ("position_sampling", C.c_int32)]


class OrSunPlane(C.Structure):
    _fields_ = [("v0", C.c_double * 3), ("v1", C.c_double * 3), ("v2", C.c_double * 3), ("v3", C.c_double * 3),
                ("distance", C.c_double), ("area", C.c_double), ("box_half_angle", C.c_double)]


class OrTraceDesc(C.Structure):
    _fields_ = [("num_rays", C.c_uint64), ("ray_begin", C.c_uint64), ("ray_count", C.c_uint64),
                ("seed", C.c_uint64), ("job_id", C.c_uint32), ("max_depth", C.c_int32), ("flags", C.c_uint32),
                ("kernel", C.c_int32), ("flux_element", C.c_int32), ("flux_nx", C.c_int32), ("flux_ny", C.c_int32),
                ("block_size", C.c_int32), ("grid_blocks", C.c_int32), ("reserved", C.c_int32),
                ("stream", C.c_void_p)]


class OrOutputs(C.Structure):
    _fields_ = [("flux", C.c_void_p), ("element_hits", C.c_void_p), ("stats", C.c_void_p),
                ("records", C.c_void_p), ("record_elements", C.c_void_p)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        d3 = C.POINTER(C.c_double)
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_philox4x32_10.restype = None
        L.or_uniform.argtypes = [C.c_uint32]
        L.or_uniform.restype = C.c_double
        L.or_halton.argtypes = [C.c_uint32, C.c_uint32]
        L.or_halton.restype = C.c_double
        L.or_stratified_cell.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32]
        L.or_stratified_cell.restype = C.c_uint32
        L.or_element_frame.argtypes = [C.c_void_p, d3]
        L.or_element_frame.restype = None
        L.or_element_ref_aabb.argtypes = [C.c_void_p, d3, d3]
        L.or_element_ref_aabb.restype = None
        L.or_scene_create.argtypes = [C.c_void_p, C.c_int32]
        L.or_scene_create.restype = C.c_void_p
        L.or_scene_destroy.argtypes = [C.c_void_p]
        L.or_scene_destroy.restype = None
        L.or_scene_set_sun.argtypes = [C.c_void_p, C.POINTER(OrSun)]
        L.or_scene_get_sun_plane.argtypes = [C.c_void_p, C.POINTER(OrSunPlane)]
        L.or_scene_get_frames.argtypes = [C.c_void_p, d3]
        L.or_closest_hit.argtypes = [C.c_void_p, d3, d3, C.c_double, C.c_double, C.c_int32, C.POINTER(OrHit)]
        L.or_intersect_element.argtypes = [C.c_void_p, C.c_int32, d3, d3, C.c_double, C.c_double, C.POINTER(OrHit)]
        L.or_sample_sun.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, d3, d3]
        L.or_sample_sun.restype = None
        L.or_perturb.argtypes = [d3, C.c_double, C.c_double, C.c_double, d3]
        L.or_perturb.restype = None
        L.or_reflect.argtypes = [d3, d3, d3]
        L.or_reflect.restype = None
        L.or_mirror_direction.argtypes = [d3, d3, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_double, d3]
        L.or_mirror_direction.restype = None
        L.or_debug_plant.argtypes = [C.c_int]
        L.or_debug_plant.restype = C.c_int
        L.or_trace.argtypes = [C.c_void_p, C.POINTER(OrTraceDesc), C.POINTER(OrOutputs), C.c_int32]
        _lib = L
    return _lib


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def philox(ctr, key):
    out = (C.c_uint32 * 4)()
    lib().or_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
    return list(out)


def uniform(x: int) -> float:
    return lib().or_uniform(x)


def halton(i: int, base: int) -> float:
    return lib().or_halton(i, base)


def stratified_cell(ray: int, seed: int = 0, job: int = 0) -> int:
    return lib().or_stratified_cell(ray, seed, job)


def _elem_ptr(elements: np.ndarray):
    arr = np.ascontiguousarray(elements)
    assert arr.dtype.itemsize == 128
    return arr, arr.ctypes.data_as(C.c_void_p)


def element_frame(element) -> np.ndarray:
    arr, p = _elem_ptr(np.atleast_1d(element))
    out = (C.c_double * 9)()
    lib().or_element_frame(p, out)
    return np.array(out).reshape(3, 3)


def element_ref_aabb(element):
    arr, p = _elem_ptr(np.atleast_1d(element))
    lo, hi = (C.c_double * 3)(), (C.c_double * 3)()
    lib().or_element_ref_aabb(p, lo, hi)
    return np.array(lo), np.array(hi)


class Scene:
    """Oracle scene: FP64 frames, its own median-split BVH, the sun plane."""

    def __init__(self, elements: np.ndarray):
        self._elements, p = _elem_ptr(elements)
        self.n = len(self._elements)
        self.h = lib().or_scene_create(p, self.n)
        if not self.h:
            raise ValueError("oracle rejected the element set")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_destroy(self.h)
            self.h = None

    def set_sun(self, vector, param, shape=0, dni=1000.0, position_sampling=0):
        s = OrSun()
        for k in range(3):
            s.vector[k] = float(vector[k])
        s.param, s.dni, s.shape, s.position_sampling = float(param), float(dni), int(shape), int(position_sampling)
        assert lib().or_scene_set_sun(self.h, C.byref(s)) == 0

    def sun_plane(self) -> dict:
        pl = OrSunPlane()
        assert lib().or_scene_get_sun_plane(self.h, C.byref(pl)) == 0
        return {"v0": np.array(pl.v0), "v1": np.array(pl.v1), "v2": np.array(pl.v2), "v3": np.array(pl.v3),
                "distance": pl.distance, "area": pl.area, "box_half_angle": pl.box_half_angle}

    def frames(self) -> np.ndarray:
        out = np.zeros((self.n, 9))
        assert lib().or_scene_get_frames(self.h, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
        return out.reshape(self.n, 3, 3)

    def closest_hit(self, o, d, tmin=1e-3, tmax=1e16, brute_force=False):
        h = OrHit()
        r = lib().or_closest_hit(self.h, _d3(o), _d3(d), tmin, tmax, int(brute_force), C.byref(h))
        if r < 0:
            raise RuntimeError("oracle traversal failed")
        return h if r else None

    def intersect(self, elem, o, d, tmin=1e-3, tmax=1e16):
        h = OrHit()
        r = lib().or_intersect_element(self.h, elem, _d3(o), _d3(d), tmin, tmax, C.byref(h))
        return h if r == 1 else None

    def sample_sun(self, ray, seed=0, job=0):
        p, d = (C.c_double * 3)(), (C.c_double * 3)()
        lib().or_sample_sun(self.h, ray, seed, job, p, d)
        return np.array(p), np.array(d)

    def trace(self, num_rays, seed=0, job_id=0, max_depth=5, flux_nx=0, flux_ny=0, flux_element=-1,
              ray_begin=0, ray_count=0, records=False, include_direct_sun=False, threads=None):
        count = ray_count or (num_rays - ray_begin)
        desc = OrTraceDesc()
        desc.num_rays, desc.ray_begin, desc.ray_count, desc.seed = num_rays, ray_begin, count, seed
        desc.job_id, desc.max_depth, desc.flux_element = job_id, max_depth, flux_element
        desc.flux_nx, desc.flux_ny = flux_nx, flux_ny
        desc.flags = (1 if records else 0) | (2 if include_direct_sun else 0)
        flux = np.zeros((max(flux_ny, 0), max(flux_nx, 0)), dtype=np.int64)
        hits = np.zeros(self.n, dtype=np.uint64)
        stats = np.zeros(STATS, dtype=np.uint64)
        rec = np.zeros((count, max_depth, 4)) if records else None
        rid = np.zeros((count, max_depth), dtype=np.uint32) if records else None
        out = OrOutputs()
        out.flux = flux.ctypes.data if flux_nx > 0 and flux_ny > 0 else None
        out.element_hits = hits.ctypes.data
        out.stats = stats.ctypes.data
        out.records = rec.ctypes.data if records else None
        out.record_elements = rid.ctypes.data if records else None
        if threads is None:
            threads = os.cpu_count() or 1
        r = lib().or_trace(self.h, C.byref(desc), C.byref(out), int(threads))
        if r != 0:
            raise RuntimeError(f"oracle trace failed ({r})")
        return {"flux": flux, "element_hits": hits, "stats": stats, "records": rec, "record_elements": rid}


def perturb(a, sigma, xi_r, xi_phi):
    out = (C.c_double * 3)()
    lib().or_perturb(_d3(a), sigma, xi_r, xi_phi, out)
    return np.array(out)


def reflect(d, n):
    out = (C.c_double * 3)()
    lib().or_reflect(_d3(d), _d3(n), out)
    return np.array(out)


def mirror_direction(d, n, sigma_slope, sigma_spec, xs_r=0.5, xs_p=0.5, xp_r=0.5, xp_p=0.5):
    out = (C.c_double * 3)()
    lib().or_mirror_direction(_d3(d), _d3(n), sigma_slope, sigma_spec, xs_r, xs_p, xp_r, xp_p, out)
    return np.array(out)


class planted:
    """Context manager: trace with planted mistake `code` switched on (tests only)."""

    def __init__(self, code: int):
        self.code = code

    def __enter__(self):
        self.prev = lib().or_debug_plant(self.code)
        return self

    def __exit__(self, *exc):
        lib().or_debug_plant(self.prev)
        return False
