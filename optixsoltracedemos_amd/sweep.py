"""Batched annual sweep over the 10k-heliostat field: 96 sun positions with
uneven ray budgets, mirrors re-aimed and the BVH refit per step.

Two ways to run the same work:
  * batched: one st_trace_batch call - FP64 GPU re-aim of every step, one
    level-wise refit over all steps, all sun planes, and ONE trace launch over
    the chunks of every step, captured in a hipGraph and replayed;
  * plain per-scene loop, the reference's update()/run() cycle
    (lib/soltrace_system.cpp:172-184, demos/demo_dynamic_scene.cpp:82-97):
    per step one set-up (new sun, re-aim, refit, sun plane and grid:
    st_scene_set_sun_and_aim), one synchronous trace, flux copied back to the
    host.
Both time the same work: the batched timer includes the copy of every step's
flux map to the host, as the plain loop copies each. Both produce identical
per-step results (same Philox keys: seed, job = step).
"""
from __future__ import annotations

import time

import numpy as np

from . import _abi, scenes, tracer


def run_batched(ctx, w, suns, budgets, seed=0, flux=(256, 128), use_graph=True, reps=2, position_sampling=_abi.POS_STRATIFIED,
                bvh_only=False):
    import torch
    sc = tracer.Scene(ctx, w.elements)
    kw = dict(seed=seed, max_depth=w.max_depth, flux_nx=flux[0], flux_ny=flux[1], use_graph=use_graph,
              bvh_only=bvh_only)
    sc.set_sun(suns[0], w.sun_param, position_sampling=position_sampling)
    out = sc.trace_batch(suns, budgets, w.extra["target"], w.sun_param, position_sampling=position_sampling, **kw)
    torch.cuda.synchronize()                                   # first call: set-up + graph capture
    host = torch.empty(out["flux"].shape, dtype=out["flux"].dtype, pin_memory=True)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = sc.trace_batch(suns, budgets, w.extra["target"], w.sun_param, position_sampling=position_sampling,
                             out=out, **kw)
        host.copy_(out["flux"])                                # every step's map on the host, as the plain loop
        times.append(time.perf_counter() - t0)
    return min(times), out, host.numpy().copy()


def run_plain_loop(ctx, w, suns, budgets, seed=0, flux=(256, 128), position_sampling=_abi.POS_STRATIFIED):
    import torch
    sc = tracer.Scene(ctx, w.elements)
    K = len(budgets)
    host = torch.empty((K, flux[1], flux[0]), dtype=torch.int64, pin_memory=True)
    desc = sc.make_desc(1, max_depth=w.max_depth, flux_nx=flux[0], flux_ny=flux[1])
    res = sc.alloc_outputs(desc)                               # device outputs reused by every step
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(K):
        sc.set_sun_and_aim(suns[k], w.sun_param, w.extra["target"], position_sampling=position_sampling)
        d = sc.make_desc(int(budgets[k]), seed=seed, job_id=k, max_depth=w.max_depth, flux_nx=flux[0],
                         flux_ny=flux[1], zero=True)
        sc.trace_into(d, res)
        host[k].copy_(res.flux)
    return time.perf_counter() - t0, host.numpy().copy()


def run_dropin_loop(w, suns, budgets, flux=(256, 128)):
    """The reference's frame loop through the C++ drop-in API (demos/demo_sweep.cpp:
    per step host FP64 re-aim of every heliostat, SolTraceSystem::update(), run(),
    flux map read back), in its own process. Returns (loop seconds, maps)."""
    import os
    import subprocess
    import tempfile
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "demo_sweep")
    with tempfile.TemporaryDirectory() as td:
        paths = [os.path.join(td, f) for f in ("elements.bin", "suns.bin", "budgets.bin", "flux.bin")]
        np.ascontiguousarray(w.elements).tofile(paths[0])
        np.ascontiguousarray(suns, dtype=np.float64).tofile(paths[1])
        np.ascontiguousarray(budgets, dtype=np.uint64).tofile(paths[2])
        r = subprocess.run([exe, *paths, str(flux[0]), str(flux[1])], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"demo_sweep failed ({r.returncode}): {r.stdout[-2000:]}")
        t = [x.strip() for x in r.stdout.strip().splitlines()[-1].split(",")]
        loop_s = float(t[t.index("loop_s") + 1])
        maps = np.fromfile(paths[3], dtype=np.int64).reshape(len(budgets), flux[1], flux[0])
    return loop_s, maps


def compare(ctx, steps=96, seed=0):
    w = scenes.heliostat_field(10_000)
    suns, budgets = scenes.annual_sweep(steps, seed=seed)
    total = int(budgets.sum())
    tb, _, fb = run_batched(ctx, w, suns, budgets, seed=seed)
    tw, _, fw = run_batched(ctx, w, suns, budgets, seed=seed, bvh_only=True)
    tp, plain = run_plain_loop(ctx, w, suns, budgets, seed=seed)
    identical = all(np.array_equal(fb[k], plain[k]) and np.array_equal(fw[k], plain[k]) for k in range(steps))
    out = {"steps": steps, "total_rays": total, "batched_s": tb, "batched_bvh_only_s": tw, "plain_loop_s": tp,
           "batched_rays_per_s": total / tb, "plain_rays_per_s": total / tp, "speedup": tp / tb,
           "identical_flux": bool(identical)}
    if seed == 0:
        td, dm = run_dropin_loop(w, suns, budgets)
        diff = float(np.abs(dm - fb).sum()) / max(float(fb.sum()), 1.0)
        out["dropin_frame_loop"] = {"loop_s": td, "rays_per_s": total / td, "batched_speedup": td / tb,
                                    "flux_l1_vs_batched": diff}
    return out
