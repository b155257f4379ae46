"""Every workload at its full size against the oracle over the same rays
(shared Philox draws): the three-heliostat scene (1e6 rays), the parabolic dish
(1e7 rays), the 10k field with Halton positions, the drop-in's default (1e8
rays; the stratified headline form is in test_gpu_parity), steps of the
batched annual sweep at their full budgets, and served jobs on scenes of 3,
100, 1,000 and 10,000 heliostats.

No records at these sizes: flux maps, per-element counts and statistics, to the
bar of test_gpu_parity: flux L1 / hits <= FLUX_L1_MAX (bin flips of
edge-grazing rays), per-element counts and statistics within the grazing-ray
fraction ID_MISMATCH_MAX of the rays. Per-ray records are compared over the
first 1e7 rays of the headline workload.
"""
import numpy as np
import pytest

from optixsoltracedemos_amd import _abi, scenes, serving, tracer
from tests.parity import flux_l1, ray_outcomes_balance
from tests.test_gpu_parity import FLUX_L1_MAX, ID_MISMATCH_MAX

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def check_full(g, o, n, what):
    l1 = flux_l1(g["flux"], o["flux"])
    print(f"{what}: rays {n:.2e} flux L1 {l1:.2e} receiver hits gpu {int(g['stats'][3])} "
          f"oracle {int(o['stats'][3])}", flush=True)
    assert l1 <= FLUX_L1_MAX, (what, l1)
    dh = np.abs(g["element_hits"].astype(np.int64) - o["element_hits"].astype(np.int64)).sum()
    assert dh <= max(4, 4 * ID_MISMATCH_MAX * n), (what, dh)
    ds = np.abs(g["stats"].astype(np.int64)[:11] - o["stats"].astype(np.int64)[:11])
    assert ds.max() <= max(2, 2 * ID_MISMATCH_MAX * n), (what, ds)
    assert ray_outcomes_balance(g["stats"]) and ray_outcomes_balance(o["stats"])


def oracle_trace(oracle_mod, elements, sun, w, n, pos, seed=0, job_id=0, flux=None, max_depth=None):
    fx, fy = flux or (w.flux_nx, w.flux_ny)
    s = oracle_mod.Scene(elements)
    s.set_sun(sun, w.sun_param, w.sun_shape, position_sampling=pos)
    return s.trace(n, seed=seed, job_id=job_id, max_depth=max_depth or w.max_depth, flux_nx=fx, flux_ny=fy,
                   threads=ORACLE_THREADS)


def reaimed(w, sun, target):
    """The workload's elements with every mirror aimed on the bisector of the
    sun and the target (what st_scene_aim_heliostats computes on the GPU)."""
    e = w.elements.copy()
    for i in range(len(e)):
        if e[i]["role"] == scenes.ROLE_MIRROR:
            e[i]["aim_point"] = scenes.bisector_aim(e[i]["origin"], target, sun)
    return e


@pytest.mark.parametrize("name,pos", [("three_heliostats", _abi.POS_PHILOX), ("parabolic_dish", _abi.POS_PHILOX),
                                      ("field_10k", _abi.POS_HALTON)])
def test_workload_full_size(gpu_ctx, oracle_mod, name, pos):
    w = scenes.WORKLOADS[name]()
    n = w.num_rays
    sc = tracer.Scene(gpu_ctx, w.elements)
    sc.set_sun(w.sun_vector, w.sun_param, w.sun_shape, position_sampling=pos)
    g = sc.trace(n, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny).numpy()
    o = oracle_trace(oracle_mod, w.elements, w.sun_vector, w, n, pos)
    check_full(g, o, n, name)
    sc.close()


def test_sweep_steps_full_size(gpu_ctx, oracle_mod):
    """The whole 96-step annual sweep in one batched call; the lowest-sun, the
    highest-sun and the largest step against the oracle on elements re-aimed for
    their suns."""
    w = scenes.heliostat_field(10_000)
    suns, budgets = scenes.annual_sweep(96, seed=0)
    tgt = w.extra["target"]
    sc = tracer.Scene(gpu_ctx, w.elements)
    sc.set_sun(suns[0], w.sun_param, position_sampling=_abi.POS_STRATIFIED)
    out = sc.trace_batch(suns, budgets, tgt, w.sun_param, seed=0, max_depth=w.max_depth, flux_nx=w.flux_nx,
                         flux_ny=w.flux_ny, position_sampling=_abi.POS_STRATIFIED)
    flux, hits, st = (out[k].cpu().numpy() for k in ("flux", "element_hits", "stats"))
    for k in sorted({int(np.argmin(suns[:, 2])), int(np.argmax(suns[:, 2])), int(np.argmax(budgets))}):
        n = int(budgets[k])
        o = oracle_trace(oracle_mod, reaimed(w, suns[k], tgt), suns[k], w, n, _abi.POS_STRATIFIED, job_id=k)
        el = np.degrees(np.arcsin(suns[k][2]))
        check_full({"flux": flux[k], "element_hits": hits[k], "stats": st[k]}, o, n, f"sweep step {k} el {el:.1f}")
    sc.close()


def test_served_jobs_full_size(gpu_ctx, oracle_mod):
    """One job per scene size, served together (the 10k job long enough for its
    own grid and reflection lists), each against the oracle."""
    mix = serving.build_serving_scenes(gpu_ctx, sizes=(3, 100, 1000, 10_000))
    srv = serving.Server(gpu_ctx, max_active_jobs=8, chunk_rays=1 << 20)
    slots = [srv.add_scene(s) for _, _, s in mix]
    suns = [scenes.sun_vector(el, az) for el, az in ((25.0, 100.0), (48.0, 210.0), (70.0, 300.0), (33.0, 160.0))]
    rays = [3_000_000, 5_000_000, 8_000_000, (1 << 24) + 777]
    flux = (64, 32)
    hs = [srv.submit(slots[i], suns[i], rays[i], seed=9, job_id=100 + i, flux_nx=flux[0], flux_ny=flux[1],
                     position_sampling=_abi.POS_STRATIFIED) for i in range(4)]
    res = [srv.wait(h) for h in hs]
    assert all(r["status"] == 0 for r in res)
    for i, h in enumerate(hs):
        n_h, w, _ = mix[i]
        fl, ht = srv.fetch(h, len(w.elements), (flux[1], flux[0]))
        target = w.elements[w.receiver]["origin"]
        o = oracle_trace(oracle_mod, reaimed(w, suns[i], target), suns[i], w, rays[i], _abi.POS_STRATIFIED,
                         seed=9, job_id=100 + i, flux=flux, max_depth=5)
        check_full({"flux": fl, "element_hits": ht, "stats": res[i]["stats"]}, o, rays[i], f"served {n_h} heliostats")
        srv.release(h)
    srv.close()


def test_field_per_ray_parity_1e7(gpu_ctx, oracle_mod):
    """Per-ray parity (element sequences, hit points, segment directions) over
    the first 1e7 rays of the headline workload (10k field, stratified
    positions, 1e8-ray job), in five slices of 2e6 rays traced by index range
    on both sides, to the bar of test_gpu_parity."""
    from tests.test_gpu_parity import check_parity
    w = scenes.heliostat_field(10_000)
    sc = tracer.Scene(gpu_ctx, w.elements)
    sc.set_sun(w.sun_vector, w.sun_param, w.sun_shape, position_sampling=_abi.POS_STRATIFIED)
    sc.build_lists()  # the headline's path: sun-plane grid, reflection and pair lists
    os_ = oracle_mod.Scene(w.elements)
    os_.set_sun(w.sun_vector, w.sun_param, w.sun_shape, position_sampling=_abi.POS_STRATIFIED)
    step = 2_000_000
    for begin in range(0, 10_000_000, step):
        kw = dict(max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny, ray_begin=begin, ray_count=step,
                  records=True)
        g = sc.trace(w.num_rays, **kw).numpy()
        o = os_.trace(w.num_rays, threads=ORACLE_THREADS, **kw)
        print(f"rays [{begin}, {begin + step}): checking", flush=True)
        check_parity(w, g, o, step)
    sc.close()
