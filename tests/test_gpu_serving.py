"""Serving loop: each job's result is the same alone and in the mix (bitwise:
draws are keyed by seed, job and ray, bins are integer atomics), HBM-footprint
admission, cancellation, and a short Poisson stream."""
import numpy as np
import pytest

from optixsoltracedemos_amd import _abi, scenes, serving, tracer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix(gpu_ctx):
    return serving.build_serving_scenes(gpu_ctx, sizes=(3, 100, 1000, 10_000))


def alone(ctx, w, sun, rays, seed, job_id, flux):
    sc = tracer.Scene(ctx, w.elements)
    target = w.elements[w.receiver]["origin"]
    sc.set_sun(sun, 4.65e-3)
    sc.aim_heliostats(sun, target)
    r = sc.trace(rays, seed=seed, job_id=job_id, max_depth=5, flux_nx=flux[0], flux_ny=flux[1]).numpy()
    return r, sc.sun_plane()["area"]


def test_results_identical_alone_and_in_the_mix(gpu_ctx, mix):
    srv = serving.Server(gpu_ctx, max_active_jobs=16, chunk_rays=1 << 18)
    slots = [srv.add_scene(s) for _, _, s in mix]
    t, kind, suns, rays = serving.poisson_jobs(24, seed=5, rays_hi=3e6)
    hs = [srv.submit(slots[kind[i]], suns[i], int(rays[i]), seed=5, job_id=i, flux_nx=32, flux_ny=16)
          for i in range(len(kind))]
    res = [srv.wait(h) for h in hs]
    assert all(r["status"] == 0 for r in res)
    assert srv.stats()["launches"] > 1
    checked = set()
    for i, h in enumerate(hs):
        if kind[i] in checked:
            continue
        checked.add(kind[i])
        n, w, _ = mix[kind[i]]
        flux, hits = srv.fetch(h, len(w.elements), (16, 32))
        ref, area = alone(gpu_ctx, w, suns[i], int(rays[i]), 5, i, (32, 16))
        assert np.array_equal(res[i]["stats"], ref["stats"]), n
        assert np.array_equal(flux, ref["flux"]), n
        assert np.array_equal(hits, ref["element_hits"]), n
        assert res[i]["power_per_ray"] == pytest.approx(1000.0 * area / int(rays[i]), rel=1e-12)
    assert len(checked) == 4
    for h in hs:
        srv.release(h)
    srv.close()


def test_long_job_with_lists_identical_alone(gpu_ctx, mix):
    """A job of 2^24 rays or more gets its own sun-plane grid and first-reflection
    lists on the server; served among short jobs it equals the job traced alone."""
    _, w, sc = mix[3]
    srv = serving.Server(gpu_ctx, max_active_jobs=8, chunk_rays=1 << 22, round_rays=6 << 20)
    slot = srv.add_scene(sc)
    sun = scenes.sun_vector(38.0, 150.0)
    rays = (1 << 24) + 12_345
    long_job = srv.submit(slot, sun, rays, seed=9, job_id=77, flux_nx=64, flux_ny=32)
    short = [srv.submit(slot, scenes.sun_vector(50.0 + k, 200.0), 300_000 + k, seed=9, job_id=k, flux_nx=64,
                        flux_ny=32) for k in range(5)]
    r = srv.wait(long_job)
    assert r["status"] == 0 and all(srv.wait(h)["status"] == 0 for h in short)
    flux, hits = srv.fetch(long_job, len(w.elements), (32, 64))
    ref, _ = alone(gpu_ctx, w, sun, rays, 9, 77, (64, 32))
    assert np.array_equal(r["stats"], ref["stats"])
    assert np.array_equal(flux, ref["flux"])
    assert np.array_equal(hits, ref["element_hits"])
    assert srv.stats()["launches"] >= 3
    for h in short + [long_job]:
        srv.release(h)
    srv.close()


def test_admission_by_hbm_footprint(gpu_ctx, mix):
    _, w, sc = mix[3]                                         # 10k field: ~7.5 MB per job
    srv = serving.Server(gpu_ctx, hbm_budget_bytes=10 << 20, max_active_jobs=8)
    slot = srv.add_scene(sc)
    hs = [srv.submit(slot, w.sun_vector, 200_000, job_id=k, flux_nx=64, flux_ny=32) for k in range(3)]
    big = srv.submit(slot, w.sun_vector, 1000, flux_nx=1024, flux_ny=1024)   # + 8 MB flux map > budget
    res = [srv.wait(h) for h in hs]
    assert all(r["status"] == 0 for r in res)
    for h in hs:
        srv.release(h)
    assert srv.wait(big)["status"] == _abi.ST_ERR_REJECTED
    st = srv.stats()
    assert st["rejected"] == 1 and st["completed"] == 3
    assert st["peak_resident_bytes"] <= 10 << 20                # never two 10k jobs at once
    srv.close()


def test_cancel(gpu_ctx, mix):
    _, w, sc = mix[3]
    srv = serving.Server(gpu_ctx, max_active_jobs=2, chunk_rays=1 << 20)
    slot = srv.add_scene(sc)
    a = srv.submit(slot, w.sun_vector, 400_000_000)
    b = srv.submit(slot, w.sun_vector, 1000, job_id=1)
    srv.cancel(a)
    ra, rb = srv.wait(a), srv.wait(b)
    assert ra["status"] in (_abi.ST_ERR_CANCELLED, 0)
    assert rb["status"] == 0
    if ra["status"] == _abi.ST_ERR_CANCELLED:
        with pytest.raises(_abi.STError):
            srv.fetch(a, len(w.elements))
    with pytest.raises(_abi.STError):
        srv.fetch(12345, 1)
    srv.release(a)
    srv.release(b)
    srv.close()


def test_short_poisson_stream(gpu_ctx, mix):
    m = serving.run_stream(gpu_ctx, n_jobs=60, rate=200.0, seed=2, rays_hi=2e6, serving_scenes=mix)
    assert m["completed"] == 60 and m["rejected"] == 0 and m["errors"] == 0
    assert 0 < m["p50_latency_ms"] <= m["p99_latency_ms"]
    assert m["aggregate_rays_per_s"] > 0


def test_served_and_plain_loop_give_identical_maps(gpu_ctx, mix):
    """The throughput comparison traces the same jobs both ways: served (all
    submitted at once) and one by one on scenes of their own."""
    for samp in (_abi.POS_STRATIFIED, _abi.POS_PHILOX):
        r = serving.compare_plain(gpu_ctx, n_jobs=24, seed=3, serving_scenes=mix, rays_hi=3e6, sampling=samp)
        assert r["identical_flux"] and r["served_s"] > 0 and r["plain_loop_s"] > 0

