"""Benchmark: traced rays/s on the 10,000-heliostat field (one MI355X).

One step = one st_trace of 1e8 sun rays over the 10k-heliostat radial-stagger
field with a 150 m tower and a cylindrical receiver (flux map 256 x 128, int64
bins, at most 4 interactions per ray). Scene, BVH and outputs are resident in
HBM before timing starts. Prints one JSON line.

  python bench.py [--gpus 1] [--steps 20] [--warmup 10] [--rays 1e8]
                  [--no-cpu-baseline] [--cpu-sample N] [--no-serve]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "traced rays/sec at 10k-heliostat field + p50/p99 job latency, 1 MI355X"
# MI355X published FP32 vector peak, 157.3 TFLOP/s, in lane operations (an FMA
# counts one): 256 CU x 128 lane-ops per clock x 2.4 GHz = 78.6e12. That
# figure counts packed FP32 (two FMAs per lane and cycle); this build issues
# plain FP32 (packed FP32 is disabled), whose rate - half of it, ~39e12 - is
# measured live by st_calibrate_valu and reported beside it.
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12
# Algorithmic lane operations per event (add/mul/fma/min/max/cmp/select/rcp/
# rsqrt/sqrt = 1, a 32x32->64 multiply = 1, sincos/sincospi = 20, atan2 = 30),
# counted from the source of each routine:
OPS_SUN_RAY = 125       # Philox4x32-10 (60), 4 uniforms (12), plane point (6), cone direction (45)
OPS_SEGMENT = 10        # hit point and local offset of a closest-hit query
OPS_NODE_VISIT = 34     # two slab tests (12 fma + 18 min/max/cmp) + order / stack
OPS_FLAT_TEST = 30      # plane t, in-rectangle test, closest-hit update
OPS_PARABOLIC_TEST = 60 # quadratic in the element frame, root choice, aperture test
OPS_CYLINDER_TEST = 70  # wall quadratic about closest approach + two caps
OPS_MIRROR_HIT = 172    # Philox (60), faceforward, slope-error cone (81), reflection (14), counts
OPS_RECEIVER_HIT = 70   # local coordinates, atan2 bin, aggregated flux add
HBM_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "trace_field10k_hbm_traffic.txt")
HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, published
RECORD_BYTES_PER_SLOT = 16 + 4   # float4 (tag, x, y, z) + u32 element id per interaction slot


def cpu_model() -> str:
    """The host CPU's model string (what lscpu prints as "Model name")."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_launches(sc, desc, res, stream, n):
    """Mean kernel time (ms, HIP events on the launch stream) of n launches after one warm-up."""
    import numpy as np
    import torch
    sc.trace_into(desc, res)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
    for i in range(n):
        starts[i].record(stream)
        sc.trace_into(desc, res)
        ends[i].record(stream)
    torch.cuda.synchronize()
    return float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)]))


def ops_per_ray(stats) -> float:
    """Algorithmic lane operations per ray from a ST_TRACE_COUNT_WORK launch."""
    s = [float(x) / float(stats[0]) for x in stats]
    return (OPS_SUN_RAY + OPS_SEGMENT * s[1] + OPS_NODE_VISIT * s[11] + OPS_FLAT_TEST * s[12]
            + OPS_PARABOLIC_TEST * s[13] + OPS_CYLINDER_TEST * s[14] + OPS_MIRROR_HIT * s[2]
            + OPS_RECEIVER_HIT * s[3])


def read_traffic():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    try:
        for line in open(HBM_TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(workload, sample: int, pos: int):
    """The oracle timed over rays [0, sample) of the workload; returns the
    baseline record and the oracle's outputs (flux, element hits, statistics)."""
    from oracle import oracle as O
    threads = max(1, min(os.cpu_count() or 1, 16))
    s = O.Scene(workload.elements)
    s.set_sun(workload.sun_vector, workload.sun_param, workload.sun_shape, position_sampling=pos)
    t0 = time.perf_counter()
    out = s.trace(workload.num_rays, ray_begin=0, ray_count=sample, max_depth=workload.max_depth,
                  flux_nx=workload.flux_nx, flux_ny=workload.flux_ny, threads=threads)
    dt = time.perf_counter() - t0
    whole = "all" if sample == workload.num_rays else "the first"
    rec = {"value": sample / dt, "unit": "rays/s", "cores": threads, "kind": "port",
           "sample": f"{whole} {sample:.0e} rays of the same {workload.num_rays:.0e}-ray field workload, "
                     f"FP64 scalar oracle, {threads} threads, {dt:.2f} s wall (~{dt * threads:.0f} thread-seconds)",
           "host_cpus": os.cpu_count(), "cpu_model": cpu_model()}
    # one thread on a smaller prefix of the same rays
    one = min(sample, 2_000_000)
    t0 = time.perf_counter()
    s.trace(workload.num_rays, ray_begin=0, ray_count=one, max_depth=workload.max_depth,
            flux_nx=workload.flux_nx, flux_ny=workload.flux_ny, threads=1)
    rec["single_thread"] = {"value": one / (time.perf_counter() - t0), "unit": "rays/s", "rays": one}
    return rec, out


def flux_accuracy(gpu, ora, rays: int) -> dict:
    """One GPU launch against the oracle over the same rays (shared Philox draws)."""
    import numpy as np
    g, o = gpu["flux"].astype(np.float64), ora["flux"].astype(np.float64)
    dh = np.abs(gpu["element_hits"].astype(np.int64) - ora["element_hits"].astype(np.int64)).sum()
    return {"rays": rays, "flux_l1_vs_oracle": float(np.abs(g - o).sum() / max(o.sum(), 1.0)),
            "receiver_hits_gpu": int(gpu["stats"][3]), "receiver_hits_oracle": int(ora["stats"][3]),
            "element_hits_l1_per_ray": float(dh) / rays}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--rays", type=float, default=1e8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=float, default=0,
                    help="rays the oracle traces for the CPU baseline (default: the whole workload, "
                         "which also gives the flux-map accuracy of one GPU launch)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-vs-plain annual sweep")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the Halton-position and record-mode launches beside the headline")
    ap.add_argument("--sampling", choices=["stratified", "philox", "halton"], default="stratified",
                    help="sun-plane position sampling (stratified: 64-ray wave-coherent cells)")
    ap.add_argument("--kernel", choices=["auto", "mega", "queued", "grouped"], default="auto",
                    help="trace loop form (auto: the tuned one; grouped: queued with list and BVH "
                         "segments in separate steps)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0,
                    help="waves per SIMD of the trace variant (0: tuned; 6 or 8: stack in private memory)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 3 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")   # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": a.numel() * 4}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (rays would shard by index "
              "range, with one integer all-reduce of the flux map). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from optixsoltracedemos_amd import scenes, tracer

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    rays = int(args.rays)
    w = scenes.heliostat_field(10_000, num_rays=rays, seed=args.seed)
    ctx = tracer.Context(0)
    sc = tracer.Scene(ctx, w.elements)
    from optixsoltracedemos_amd import _abi
    pos = {"stratified": _abi.POS_STRATIFIED, "philox": _abi.POS_PHILOX, "halton": _abi.POS_HALTON}[args.sampling]
    sc.set_sun(w.sun_vector, w.sun_param, w.sun_shape, position_sampling=pos)
    tuning = None
    if not args.no_autotune and args.block == 0 and args.grid == 0 and args.waves == 0:
        tuning = sc.autotune(sample_rays=rays, max_depth=w.max_depth)   # tune at the measured size
    stream = torch.cuda.Stream(device=0)
    kernel = {"auto": _abi.KERNEL_AUTO, "mega": _abi.KERNEL_MEGA, "queued": _abi.KERNEL_QUEUED,
              "grouped": _abi.KERNEL_QUEUED_GROUPED}[args.kernel]
    desc = sc.make_desc(rays, seed=args.seed, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny,
                        zero=False, block_size=args.block, grid_blocks=args.grid, waves_per_simd=args.waves,
                        stream=stream.cuda_stream,
                        kernel=kernel)
    res = sc.alloc_outputs(desc)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        sc.trace_into(desc, res)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        sc.trace_into(desc, res)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(starts, ends)]))
    stats = res.stats.cpu().numpy().astype(np.int64)
    calls = args.warmup + args.steps
    assert stats[0] == calls * rays, "every launched ray must be counted once"
    per_call = stats / calls
    value = args.steps * rays / wall
    # one more launch of the same work, untimed, counting node visits and tests
    wdesc = sc.make_desc(rays, seed=args.seed, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny,
                         zero=True, block_size=args.block, grid_blocks=args.grid, waves_per_simd=args.waves,
                         stream=stream.cuda_stream,
                         kernel=kernel, count_work=True)
    wres = sc.alloc_outputs(wdesc)
    sc.trace_into(wdesc, wres)
    torch.cuda.synchronize()
    work = wres.stats.cpu().numpy().astype(np.int64)
    assert np.array_equal(work[:11], stats[:11] // calls), "work-count launch must trace the same rays"
    w_ops = ops_per_ray(work)
    achieved = w_ops * rays / (kernel_ms * 1e-3) / 1e12
    plain_peak = ctx.calibrate_valu() / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "field_10k_heliostats_cylinder_receiver", "heliostats": 10_000,
                   "rays_per_step": rays, "flux_bins": f"{w.flux_nx}x{w.flux_ny}", "max_depth": w.max_depth,
                   "sun": "pillbox 4.65 mrad, el 60 az 180", "slope_error_mrad": 1.5, "reflectivity": 0.93,
                   "seed": args.seed, "sampling": args.sampling, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tlane-op/s",
                     "frac": achieved / VALU_PEAK_TOPS, "traffic": read_traffic(),
                     "peak_plain_fp32_measured": plain_peak, "frac_of_plain_fp32": achieved / plain_peak,
                     "ops_per_ray": w_ops, "kernel_ms": kernel_ms,
                     "work_per_ray": {"segments": float(work[1]) / rays, "node_visits": float(work[11]) / rays,
                                      "flat_tests": float(work[12]) / rays,
                                      "parabolic_tests": float(work[13]) / rays,
                                      "cylinder_tests": float(work[14]) / rays}},
        "tuning": tuning,
        "segments_per_s": float(per_call[1]) * args.steps / wall,
        "receiver_hits_per_ray": float(per_call[3] / rays),
    }
    line["launch"] = sc.last_launch()
    if not args.no_variants:
        # the reference's raygen (Halton positions, cuda/sun.cu:16-38) on the same rays
        sc.set_sun(w.sun_vector, w.sun_param, w.sun_shape, position_sampling=_abi.POS_HALTON)
        sc.build_lists()   # as the headline's tuner does: no lazy list build inside the timed launches
        hdesc = sc.make_desc(rays, seed=args.seed, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny,
                             zero=True, block_size=args.block, grid_blocks=args.grid, waves_per_simd=args.waves,
                             stream=stream.cuda_stream,
                             kernel=kernel)
        hres = sc.alloc_outputs(hdesc)
        hms = time_launches(sc, hdesc, hres, stream, 5)
        line["halton_positions"] = {"kernel_ms": hms, "rays_per_s": rays / (hms * 1e-3), "launch": sc.last_launch(),
                                    "receiver_hits_per_ray": float(hres.stats.cpu()[3]) / rays}
        # reference-compatible record mode: every launch writes the hit-record
        # buffer (lib/soltrace_system.cpp:118-125) plus the element id of each slot
        sc.set_sun(w.sun_vector, w.sun_param, w.sun_shape, position_sampling=pos)
        sc.build_lists()
        rdesc = sc.make_desc(rays, seed=args.seed, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny,
                             zero=True, records=True, block_size=args.block, grid_blocks=args.grid, waves_per_simd=args.waves,
                             stream=stream.cuda_stream, kernel=kernel)
        rres = sc.alloc_outputs(rdesc)
        rms = time_launches(sc, rdesc, rres, stream, 3)
        rbytes = float(rays) * w.max_depth * RECORD_BYTES_PER_SLOT
        line["record_mode"] = {"kernel_ms": rms, "rays_per_s": rays / (rms * 1e-3),
                               "record_bytes_per_launch": rbytes,
                               "record_gbs": rbytes / (rms * 1e-3) / 1e9,
                               "frac_of_hbm_peak": rbytes / (rms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               "hbm_floor_ms": rbytes / (HBM_PEAK_GBS * 1e9) * 1e3}
        del rres
        torch.cuda.empty_cache()
    if not args.no_serve:
        try:
            from optixsoltracedemos_amd import serving
            scs = serving.build_serving_scenes(ctx)
            line["latency"] = serving.run_stream(ctx, n_jobs=2000, rate=200.0, seed=args.seed, serving_scenes=scs)
            line["serving_load"] = [serving.run_stream(ctx, n_jobs=2000, rate=r, seed=args.seed, serving_scenes=scs)
                                    for r in (500.0, 1000.0, 2000.0, 5000.0)]
            line["served_vs_plain"] = serving.compare_plain(ctx, n_jobs=400, seed=args.seed, serving_scenes=scs)
        except Exception as ex:  # reported, never silently replaced
            line["latency"] = {"error": f"{type(ex).__name__}: {ex}"}
    if not args.no_sweep:
        try:
            from optixsoltracedemos_amd import sweep
            line["sweep"] = sweep.compare(ctx, steps=96, seed=args.seed)
        except Exception as ex:  # reported, never silently replaced
            line["sweep"] = {"error": f"{type(ex).__name__}: {ex}"}
    if not args.no_cpu_baseline:
        sample = int(args.cpu_sample) or rays
        line["cpu_baseline"], ora = cpu_baseline(w, sample, pos)
        if sample == rays:
            line["accuracy"] = flux_accuracy(wres.numpy(), ora, rays)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
