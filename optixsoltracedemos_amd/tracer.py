"""Python host layer over the C ABI (include/soltrace_hip.h).

Device buffers are torch tensors on ``cuda:<device>`` (plumbing only); every
computation runs in the HIP library. Nothing here falls back to the CPU: if the
library or the GPU is missing the calls raise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import Outputs, SunPlane, TraceDesc, call, load


def _torch():
    import torch  # imported lazily: only device memory and streams come from torch
    return torch


class Context:
    def __init__(self, device: int = 0):
        load()
        h = C.c_void_p()
        call("st_context_create", int(device), C.byref(h))
        self.h = h
        self.device = int(device)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            call("st_context_destroy", self.h)
            self.h = C.c_void_p()

    def calibrate_valu(self) -> float:
        """Measured plain FP32 vector issue rate of the device, lane operations per second."""
        v = C.c_double()
        call("st_calibrate_valu", self.h, C.byref(v))
        return v.value

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class TraceResult:
    flux: object | None
    element_hits: object | None
    stats: object | None
    records: object | None
    record_elements: object | None

    def numpy(self):
        f = lambda t: None if t is None else t.cpu().numpy()
        return {"flux": f(self.flux), "element_hits": f(self.element_hits), "stats": f(self.stats),
                "records": f(self.records), "record_elements": f(self.record_elements)}


class Scene:
    def __init__(self, ctx: Context, elements: np.ndarray):
        self.ctx = ctx
        carr, n = _abi.elements_to_ctypes(elements)
        h = C.c_void_p()
        call("st_scene_create", ctx.h, carr, n, C.byref(h))
        self.h = h
        self.n = n

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            call("st_scene_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scene state ---------------------------------------------------------
    def set_sun(self, vector, param, shape=_abi.SUN_PILLBOX, dni=1000.0, position_sampling=_abi.POS_PHILOX):
        s = _abi.sun_desc(vector, param, shape, dni, position_sampling)
        call("st_scene_set_sun", self.h, C.byref(s))

    def sun_plane(self) -> dict:
        pl = SunPlane()
        call("st_scene_get_sun_plane", self.h, C.byref(pl))
        return {"v0": np.array(pl.v0), "v1": np.array(pl.v1), "v2": np.array(pl.v2), "v3": np.array(pl.v3),
                "distance": pl.distance, "area": pl.area, "box_half_angle": pl.box_half_angle}

    def frames(self) -> np.ndarray:
        out = np.zeros((self.n, 9))
        call("st_scene_get_frames", self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.n, 3, 3)

    def aabbs(self) -> np.ndarray:
        out = np.zeros((self.n, 6), dtype=np.float32)
        call("st_scene_get_aabbs", self.h, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def bvh_info(self) -> dict:
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        call("st_scene_bvh_info", self.h, C.byref(a), C.byref(b), C.byref(c))
        return {"nodes": a.value, "depth": b.value, "max_leaf": c.value}

    def update_aim(self, aim_points: np.ndarray, zrot_deg: np.ndarray | None = None):
        a = np.ascontiguousarray(aim_points, dtype=np.float64).reshape(self.n, 3)
        z = None if zrot_deg is None else np.ascontiguousarray(zrot_deg, dtype=np.float64)
        call("st_scene_update_aim", self.h, a.ctypes.data_as(C.POINTER(C.c_double)),
             None if z is None else z.ctypes.data_as(C.POINTER(C.c_double)))

    def update_elements(self, elements: np.ndarray):
        carr, n = _abi.elements_to_ctypes(elements)
        call("st_scene_update_elements", self.h, carr, n)

    def set_sun_and_aim(self, vector, param, target, shape=_abi.SUN_PILLBOX, dni=1000.0,
                        position_sampling=_abi.POS_PHILOX):
        """set_sun + aim_heliostats(vector, target) with one sun-plane / grid build."""
        s = _abi.sun_desc(vector, param, shape, dni, position_sampling)
        t = (C.c_double * 3)(*[float(x) for x in target])
        call("st_scene_set_sun_and_aim", self.h, C.byref(s), t)

    def update(self, elements: np.ndarray | None, vector=None, param=4.65e-3, shape=_abi.SUN_PILLBOX, dni=1000.0,
               position_sampling=_abi.POS_PHILOX):
        """st_scene_update: new element poses and/or a new sun in one set-up."""
        carr, n = _abi.elements_to_ctypes(elements) if elements is not None else (None, 0)
        s = _abi.sun_desc(vector, param, shape, dni, position_sampling) if vector is not None else None
        call("st_scene_update", self.h, carr, n, C.byref(s) if s is not None else None)

    def aim_heliostats(self, sun_vector, target):
        s = (C.c_double * 3)(*[float(x) for x in sun_vector])
        t = (C.c_double * 3)(*[float(x) for x in target])
        call("st_scene_aim_heliostats", self.h, s, t)

    def autotune(self, sample_rays: int = 20_000_000, max_depth: int = 5) -> dict:
        r = _abi.TuneResult()
        call("st_scene_autotune", self.h, int(sample_rays), int(max_depth), C.byref(r))
        return {k: getattr(r, k) for k, _ in _abi.TuneResult._fields_ if k != "reserved"}

    def build_lists(self):
        call("st_scene_build_lists", self.h)

    def list_info(self) -> dict:
        g, r, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
        call("st_scene_list_info", self.h, C.byref(g), C.byref(r), C.byref(p))
        return {"grid_items": g.value, "reflection_items": r.value, "pair_items": p.value}

    def synchronize(self):
        call("st_scene_synchronize", self.h)

    # -- tracing -------------------------------------------------------------
    def make_desc(self, num_rays, seed=0, job_id=0, max_depth=5, flux_nx=0, flux_ny=0, flux_element=-1,
                  ray_begin=0, ray_count=0, records=False, include_direct_sun=False, zero=True,
                  kernel=_abi.KERNEL_AUTO, block_size=0, grid_blocks=0, stream=None, bvh_only=False,
                  count_work=False, waves_per_simd=0) -> TraceDesc:
        d = TraceDesc()
        d.num_rays, d.ray_begin, d.ray_count, d.seed = int(num_rays), int(ray_begin), int(ray_count), int(seed)
        d.job_id, d.max_depth = int(job_id), int(max_depth)
        d.flags = ((_abi.TRACE_RECORDS if records else 0) | (_abi.TRACE_FLUX_DIRECT_SUN if include_direct_sun else 0)
                   | (_abi.TRACE_ZERO_OUTPUTS if zero else 0) | (_abi.TRACE_BVH_ONLY if bvh_only else 0)
                   | (_abi.TRACE_COUNT_WORK if count_work else 0))
        d.kernel, d.flux_element, d.flux_nx, d.flux_ny = int(kernel), int(flux_element), int(flux_nx), int(flux_ny)
        d.block_size, d.grid_blocks, d.waves_per_simd = int(block_size), int(grid_blocks), int(waves_per_simd)
        d.stream = stream
        return d

    def alloc_outputs(self, desc: TraceDesc, element_hits=True, stats=True) -> TraceResult:
        torch = _torch()
        dev = f"cuda:{self.ctx.device}"
        count = desc.ray_count or (desc.num_rays - desc.ray_begin)
        flux = torch.zeros((desc.flux_ny, desc.flux_nx), dtype=torch.int64, device=dev) \
            if desc.flux_nx > 0 and desc.flux_ny > 0 else None
        hits = torch.zeros(self.n, dtype=torch.int64, device=dev) if element_hits else None
        st = torch.zeros(_abi.STAT_COUNT, dtype=torch.int64, device=dev) if stats else None
        rec = rid = None
        if desc.flags & _abi.TRACE_RECORDS:
            rec = torch.empty((count, desc.max_depth, 4), dtype=torch.float32, device=dev)
            rid = torch.empty((count, desc.max_depth), dtype=torch.int32, device=dev)
        return TraceResult(flux, hits, st, rec, rid)

    @staticmethod
    def outputs_struct(res: TraceResult) -> Outputs:
        o = Outputs()
        o.flux = res.flux.data_ptr() if res.flux is not None else None
        o.element_hits = res.element_hits.data_ptr() if res.element_hits is not None else None
        o.stats = res.stats.data_ptr() if res.stats is not None else None
        o.records = res.records.data_ptr() if res.records is not None else None
        o.record_elements = res.record_elements.data_ptr() if res.record_elements is not None else None
        return o

    def trace_into(self, desc: TraceDesc, res: TraceResult):
        o = self.outputs_struct(res)
        call("st_trace", self.h, C.byref(desc), C.byref(o))

    def trace_noalloc_into(self, desc: TraceDesc, res: TraceResult):
        """st_trace_noalloc: the same trace with no device allocation in the call."""
        o = self.outputs_struct(res)
        call("st_trace_noalloc", self.h, C.byref(desc), C.byref(o))

    def trace(self, num_rays, **kw) -> TraceResult:
        desc = self.make_desc(num_rays, **kw)
        res = self.alloc_outputs(desc)
        self.trace_into(desc, res)
        return res

    def last_launch(self) -> dict:
        """The trace-kernel variant the scene's last trace launched."""
        li = _abi.LaunchInfo()
        call("st_scene_last_launch", self.h, C.byref(li))
        return {k: getattr(li, k) for k, _ in _abi.LaunchInfo._fields_}

    def flux_to_watts(self, desc: TraceDesc, flux):
        torch = _torch()
        out = torch.zeros(flux.shape, dtype=torch.float64, device=flux.device)
        ppr = C.c_double()
        call("st_flux_to_watts", self.h, C.byref(desc), flux.data_ptr(), out.data_ptr(), C.byref(ppr))
        return out, ppr.value

    def trace_batch(self, sun_vectors: np.ndarray, budgets: np.ndarray, target, sun_param, sun_shape=_abi.SUN_PILLBOX,
                    seed=0, max_depth=5, flux_nx=0, flux_ny=0, flux_element=-1, use_graph=True, dni=1000.0,
                    include_direct_sun=False, out=None, stream=None, position_sampling=_abi.POS_PHILOX,
                    bvh_only=False, count_work=False):
        torch = _torch()
        K = len(budgets)
        suns = np.ascontiguousarray(sun_vectors, dtype=np.float64).reshape(K, 3)
        bud = np.ascontiguousarray(budgets, dtype=np.uint64)
        d = _abi.BatchDesc()
        d.steps, d.max_depth = K, int(max_depth)
        d.sun_vectors = suns.ctypes.data
        d.budgets = bud.ctypes.data
        for k in range(3):
            d.target[k] = float(target[k])
        d.sun = _abi.sun_desc([0, 0, 1], sun_param, sun_shape, dni, position_sampling)
        d.seed = int(seed)
        d.flux_element, d.flux_nx, d.flux_ny = int(flux_element), int(flux_nx), int(flux_ny)
        d.flags = (_abi.TRACE_FLUX_DIRECT_SUN if include_direct_sun else 0) | (_abi.TRACE_BVH_ONLY if bvh_only else 0) \
            | (_abi.TRACE_COUNT_WORK if count_work else 0)
        d.use_graph = 1 if use_graph else 0
        d.stream = stream
        dev = f"cuda:{self.ctx.device}"
        if out is None:
            out = {
                "flux": torch.zeros((K, max(flux_ny, 0), max(flux_nx, 0)), dtype=torch.int64, device=dev),
                "element_hits": torch.zeros((K, self.n), dtype=torch.int64, device=dev),
                "stats": torch.zeros((K, _abi.STAT_COUNT), dtype=torch.int64, device=dev),
            }
        fl = out["flux"].data_ptr() if flux_nx > 0 and flux_ny > 0 else None
        call("st_trace_batch", self.h, C.byref(d), fl, out["element_hits"].data_ptr(), out["stats"].data_ptr())
        return out


def stats_dict(stats) -> dict:
    a = np.asarray(stats).astype(np.int64)
    return {name: int(a[i]) for i, name in enumerate(_abi.STAT_NAMES)}
