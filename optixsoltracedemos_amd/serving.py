"""Python host layer of the continuous-batching trace service (st_server_*)
and the Poisson serving workload (device-resident scenes of 3 / 100 / 1,000 /
10,000 heliostats, random suns, log-uniform ray budgets)."""
from __future__ import annotations

import ctypes as C
import math
import threading
import time

import numpy as np

from . import _abi, scenes, tracer
from ._abi import JobDesc, JobResult, ServerConfig, ServerStats, call

ST_JOB_KEEP_AIM = 0x100


class Server:
    def __init__(self, ctx: tracer.Context, hbm_budget_bytes: int = 64 << 30, max_active_jobs: int = 64,
                 chunk_rays: int = 8 << 20, block_size: int = 0, grid_blocks: int = 0, round_rays: int = 0):
        cfg = ServerConfig()
        cfg.hbm_budget_bytes, cfg.max_active_jobs, cfg.chunk_rays = int(hbm_budget_bytes), int(max_active_jobs), int(chunk_rays)
        cfg.block_size, cfg.grid_blocks, cfg.round_rays = int(block_size), int(grid_blocks), int(round_rays)
        h = C.c_void_p()
        call("st_server_create", ctx.h, C.byref(cfg), C.byref(h))
        self.h, self.ctx, self.scenes = h, ctx, []

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            call("st_server_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene: tracer.Scene) -> int:
        slot = C.c_int32()
        call("st_server_add_scene", self.h, scene.h, C.byref(slot))
        self.scenes.append(scene)
        return slot.value

    def submit(self, slot, sun_vector, num_rays, seed=0, job_id=0, max_depth=5, flux_nx=0, flux_ny=0,
               sun_param=4.65e-3, sun_shape=_abi.SUN_PILLBOX, keep_aim=False, include_direct_sun=False,
               position_sampling=_abi.POS_PHILOX) -> int:
        j = JobDesc()
        j.scene_slot = int(slot)
        j.sun = _abi.sun_desc(sun_vector, sun_param, sun_shape, 1000.0, position_sampling)
        j.num_rays, j.seed, j.job_id, j.max_depth = int(num_rays), int(seed), int(job_id), int(max_depth)
        j.flux_nx, j.flux_ny = int(flux_nx), int(flux_ny)
        j.flags = (ST_JOB_KEEP_AIM if keep_aim else 0) | (_abi.TRACE_FLUX_DIRECT_SUN if include_direct_sun else 0)
        h = C.c_int64()
        call("st_server_submit", self.h, C.byref(j), C.byref(h))
        return h.value

    @staticmethod
    def _result(r: JobResult) -> dict:
        return {"status": r.status, "done": bool(r.done), "submit_s": r.submit_s, "start_s": r.start_s,
                "finish_s": r.finish_s, "power_per_ray": r.power_per_ray, "stats": np.array(r.stats[:], dtype=np.uint64)}

    def poll(self, h) -> dict:
        r = JobResult()
        call("st_server_poll", self.h, C.c_int64(h), C.byref(r))
        return self._result(r)

    def wait(self, h) -> dict:
        r = JobResult()
        call("st_server_wait", self.h, C.c_int64(h), C.byref(r))
        return self._result(r)

    def fetch(self, h, n_elements, flux_shape=None):
        flux = np.zeros(flux_shape, dtype=np.int64) if flux_shape else None
        hits = np.zeros(n_elements, dtype=np.uint64)
        call("st_server_fetch", self.h, C.c_int64(h), flux.ctypes.data if flux is not None else None, hits.ctypes.data)
        return flux, hits

    def cancel(self, h):
        call("st_server_cancel", self.h, C.c_int64(h))

    def release(self, h):
        call("st_server_release", self.h, C.c_int64(h))

    def stats(self) -> dict:
        s = ServerStats()
        call("st_server_get_stats", self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}


def build_serving_scenes(ctx, sizes=(3, 100, 1000, 10_000), seed=0):
    out = []
    for n in sizes:
        w = scenes.three_heliostats() if n == 3 else scenes.heliostat_field(n, seed=seed)
        sc = tracer.Scene(ctx, w.elements)
        sc.set_sun(w.sun_vector, w.sun_param)
        out.append((n, w, sc))
    return out


def poisson_jobs(n_jobs=2000, rate=200.0, seed=0, weights=(0.4, 0.3, 0.2, 0.1), rays_lo=1e5, rays_hi=1e8):
    """Arrival times, scene index, sun vector and ray budget of every job (Philox stream of `seed`)."""
    rng = np.random.Generator(np.random.Philox(seed + 17))
    t = np.cumsum(rng.exponential(1.0 / rate, n_jobs))
    kind = rng.choice(len(weights), size=n_jobs, p=np.asarray(weights) / np.sum(weights))
    el = np.degrees(np.arcsin(rng.uniform(math.sin(math.radians(5)), 1.0, n_jobs)))   # el > 5 deg
    az = rng.uniform(0.0, 360.0, n_jobs)
    rays = np.exp(rng.uniform(math.log(rays_lo), math.log(rays_hi), n_jobs)).astype(np.int64)
    suns = np.stack([scenes.sun_vector(e, a) for e, a in zip(el, az)])
    return t, kind, suns, rays


def run_stream(ctx, n_jobs=2000, rate=200.0, seed=0, hbm_budget_bytes=64 << 30, max_active_jobs=64,
               chunk_rays=8 << 20, flux=(64, 32), rays_hi=1e8, serving_scenes=None, sampling=_abi.POS_STRATIFIED):
    """Submit a Poisson stream at its arrival times, wait for all, report latency and throughput.
    Jobs sample sun positions stratified (wave-coherent 64-ray groups) unless `sampling` says otherwise."""
    sc_list = serving_scenes or build_serving_scenes(ctx)
    srv = Server(ctx, hbm_budget_bytes, max_active_jobs, chunk_rays)
    slots = [srv.add_scene(s) for _, _, s in sc_list]
    t, kind, suns, rays = poisson_jobs(n_jobs, rate, seed, rays_hi=rays_hi)
    handles = [None] * n_jobs
    errors = []

    def submitter():
        t0 = time.perf_counter()
        for i in range(n_jobs):
            dt = t[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            try:
                handles[i] = srv.submit(slots[kind[i]], suns[i], int(rays[i]), seed=seed, job_id=i,
                                        flux_nx=flux[0], flux_ny=flux[1], position_sampling=sampling)
            except Exception as ex:  # recorded, the job counts as failed
                errors.append((i, str(ex)))

    th = threading.Thread(target=submitter)
    th.start()
    results = [None] * n_jobs
    for i in range(n_jobs):
        while handles[i] is None and th.is_alive():
            time.sleep(1e-4)
        if handles[i] is None:
            continue
        results[i] = srv.wait(handles[i])
        srv.release(handles[i])
    th.join()
    st = srv.stats()
    srv.close()
    ok = [(i, r) for i, r in enumerate(results) if r is not None and r["status"] == 0]
    lat = np.array([r["finish_s"] - r["submit_s"] for _, r in ok])
    traced = int(sum(int(rays[i]) for i, _ in ok))
    span = max(r["finish_s"] for _, r in ok) - min(r["submit_s"] for _, r in ok) if ok else float("nan")
    return {
        "jobs": n_jobs, "admitted": int(st["admitted"]), "completed": len(ok), "rejected": int(st["rejected"]),
        "errors": len(errors),
        "arrival_rate_per_s": rate, "offered_rays_per_s": float(rays.sum() / t[-1]),
        "p50_latency_ms": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
        "p99_latency_ms": float(np.percentile(lat, 99) * 1e3) if len(lat) else None,
        "mean_latency_ms": float(lat.mean() * 1e3) if len(lat) else None,
        "aggregate_rays_per_s": traced / span if ok else None,
        "served_vs_offered": (traced / span) / float(rays.sum() / t[-1]) if ok else None,
        "launches": int(st["launches"]), "peak_resident_mb": st["peak_resident_bytes"] / 2**20,
        "gpu_busy_s": st["busy_s"], "span_s": span,
        "sampling": {_abi.POS_PHILOX: "philox", _abi.POS_HALTON: "halton", _abi.POS_STRATIFIED: "stratified"}[sampling],
    }


def quick_latency(ctx, n_jobs=400, seed=0, serving_scenes=None):
    """Short horizon of the serving stream for bench.py (~2 s of arrivals at 200 jobs/s)."""
    return run_stream(ctx, n_jobs=n_jobs, seed=seed, serving_scenes=serving_scenes)


def served_mix(ctx, n_jobs=200, seed=0, flux=(64, 32), max_active_jobs=64, serving_scenes=None,
               weights=(0.4, 0.3, 0.2, 0.1), sampling=_abi.POS_STRATIFIED, rays_hi=1e8, **server_kw):
    """The served half of compare_plain: n_jobs jobs of the serving mix, all
    submitted at once to one server, then waited for and their maps fetched.
    Returns (maps, seconds, server stats)."""
    import torch
    sc_list = serving_scenes or build_serving_scenes(ctx)
    _, kind, suns, rays = poisson_jobs(n_jobs, seed=seed, weights=weights, rays_hi=rays_hi)
    srv = Server(ctx, max_active_jobs=max_active_jobs, **server_kw)
    slots = [srv.add_scene(s) for _, _, s in sc_list]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hs = [srv.submit(slots[kind[i]], suns[i], int(rays[i]), seed=seed, job_id=i, flux_nx=flux[0], flux_ny=flux[1],
                     position_sampling=sampling) for i in range(n_jobs)]
    maps = []
    for i, h in enumerate(hs):
        r = srv.wait(h)
        if r["status"] != 0:
            raise RuntimeError(f"job {i} failed: {r['status']}")
        n_el = len(sc_list[kind[i]][1].elements)
        maps.append(srv.fetch(h, n_el, (flux[1], flux[0]))[0])
        srv.release(h)
    seconds = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return maps, seconds, st


def plain_loop(ctx, n_jobs=200, seed=0, flux=(64, 32), serving_scenes=None, weights=(0.4, 0.3, 0.2, 0.1),
               sampling=_abi.POS_STRATIFIED, rays_hi=1e8):
    """The plain half of compare_plain: per job, on a device-resident scene of
    its kind, set the sun and re-aim the field in one set-up
    (st_scene_set_sun_and_aim), trace synchronously and copy the flux map to
    the host (what a caller of lib/soltrace_system.cpp's update()/run() does per
    scene state). Returns (maps, seconds)."""
    import torch
    sc_list = serving_scenes or build_serving_scenes(ctx)
    _, kind, suns, rays = poisson_jobs(n_jobs, seed=seed, weights=weights, rays_hi=rays_hi)
    plain_scenes = []
    for _, w, _ in sc_list:
        ps = tracer.Scene(ctx, w.elements)
        ps.set_sun(w.sun_vector, w.sun_param)
        res = ps.alloc_outputs(ps.make_desc(1, max_depth=5, flux_nx=flux[0], flux_ny=flux[1]))  # reused per job
        plain_scenes.append((w, ps, res))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    maps = []
    for i in range(n_jobs):
        w, ps, res = plain_scenes[kind[i]]
        ps.set_sun_and_aim(suns[i], 4.65e-3, w.elements[w.receiver]["origin"], position_sampling=sampling)
        ps.trace_into(ps.make_desc(int(rays[i]), seed=seed, job_id=i, max_depth=5, flux_nx=flux[0], flux_ny=flux[1],
                                   zero=True), res)
        maps.append(res.flux.cpu().numpy())
    return maps, time.perf_counter() - t0


def compare_plain(ctx, n_jobs=200, seed=0, flux=(64, 32), max_active_jobs=64, serving_scenes=None, plain=True,
                  weights=(0.4, 0.3, 0.2, 0.1), sampling=_abi.POS_STRATIFIED, rays_hi=1e8, **server_kw):
    """Aggregate throughput of the served form against the plain per-scene loop
    on the same jobs of the serving mix (all submitted at once: throughput, not
    latency): served_mix, then plain_loop (skipped when plain is False). The
    served timer includes fetching every map to the host too. Both give every
    job's flux map bitwise the same."""
    sc_list = serving_scenes or build_serving_scenes(ctx)
    served, t_served, st = served_mix(ctx, n_jobs, seed, flux, max_active_jobs, sc_list, weights, sampling, rays_hi,
                                      **server_kw)
    if not plain:
        return {"jobs": n_jobs, "served_s": t_served, "launches": int(st["launches"]), "gpu_busy_s": st["busy_s"]}
    plain_maps, t_plain = plain_loop(ctx, n_jobs, seed, flux, sc_list, weights, sampling, rays_hi)
    same = all(np.array_equal(a, b) for a, b in zip(served, plain_maps))
    total = int(poisson_jobs(n_jobs, seed=seed, weights=weights, rays_hi=rays_hi)[3].sum())
    return {"jobs": n_jobs, "total_rays": total, "served_s": t_served, "plain_loop_s": t_plain,
            "max_active_jobs": max_active_jobs,
            "served_rays_per_s": total / t_served, "plain_rays_per_s": total / t_plain,
            "speedup": t_plain / t_served, "identical_flux": bool(same)}

