"""Oracle pins for sampling and for whole traces: distribution tests (KS) of the
sun cone and error models, closed-form illumination and power, conservation of
ray outcomes, reflectivity, depth-cap semantics, thread independence."""
import math

import numpy as np
import pytest
from scipy import stats as sst

from optixsoltracedemos_amd import scenes
from optixsoltracedemos_amd._abi import ROLE_MIRROR, ROLE_RECEIVER, SUN_GAUSSIAN, SURFACE_FLAT, POS_HALTON
from tests.parity import ray_outcomes_balance


def test_pillbox_cone_distribution(oracle_mod):
    """cuda/sun.cu:72-91: directions uniform over the cap of half-angle a around -s:
    cos(theta) uniform on [cos a, 1], azimuth uniform."""
    w = scenes.three_heliostats()
    s = oracle_mod.Scene(w.elements)
    a = 0.05
    s.set_sun(w.sun_vector, a)
    wv = -w.sun_vector / np.linalg.norm(w.sun_vector)
    ct, ph = [], []
    u = np.cross([1.0, 0, 0], wv)
    u /= np.linalg.norm(u)
    v = np.cross(wv, u)
    for i in range(4000):
        _, d = s.sample_sun(i, seed=3)
        assert abs(np.linalg.norm(d) - 1) < 1e-14
        c = d @ wv
        assert c >= math.cos(a) - 1e-15
        ct.append(c)
        ph.append(math.atan2(d @ v, d @ u) % (2 * math.pi))
    assert sst.kstest(ct, sst.uniform(loc=math.cos(a), scale=1 - math.cos(a)).cdf).pvalue > 1e-3
    assert sst.kstest(ph, sst.uniform(loc=0, scale=2 * math.pi).cdf).pvalue > 1e-3


def test_sun_positions_uniform_on_plane(oracle_mod):
    w = scenes.three_heliostats()
    s = oracle_mod.Scene(w.elements)
    s.set_sun(w.sun_vector, w.sun_param)
    pl = s.sun_plane()
    e1, e2 = pl["v1"] - pl["v0"], pl["v3"] - pl["v0"]
    ab = []
    for i in range(3000):
        p, _ = s.sample_sun(i, seed=9)
        r = p - pl["v0"]
        ab.append((r @ e1 / (e1 @ e1), r @ e2 / (e2 @ e2)))
        assert abs(r @ np.cross(e1, e2)) < 1e-9   # on the plane
    ab = np.array(ab)
    assert ab.min() > 0 and ab.max() < 1
    for k in range(2):
        assert sst.kstest(ab[:, k], "uniform").pvalue > 1e-3


def test_halton_positions(oracle_mod):
    w = scenes.three_heliostats()
    s = oracle_mod.Scene(w.elements)
    s.set_sun(w.sun_vector, w.sun_param, position_sampling=POS_HALTON)
    pl = s.sun_plane()
    e1, e2 = pl["v1"] - pl["v0"], pl["v3"] - pl["v0"]
    for i in (0, 1, 2, 3, 4, 77):
        p, _ = s.sample_sun(i)
        r = p - pl["v0"]
        assert r @ e1 / (e1 @ e1) == pytest.approx(oracle_mod.halton(i, 2), abs=1e-12)
        assert r @ e2 / (e2 @ e2) == pytest.approx(oracle_mod.halton(i, 3), abs=1e-12)


def test_gaussian_perturbation_is_rayleigh(oracle_mod):
    """Isotropic 2-D Gaussian with per-axis sigma: theta^2 / (2 sigma^2) ~ Exp(1),
    and each tangent component ~ N(0, sigma^2)."""
    rng = np.random.default_rng(2)
    a = np.array([0.2, -0.3, 0.9])
    a /= np.linalg.norm(a)
    sig = 2e-3
    u = np.cross([1.0, 0, 0], a)
    u /= np.linalg.norm(u)
    v = np.cross(a, u)
    th2, cu = [], []
    for xr, xp in rng.uniform(1e-7, 1, size=(4000, 2)):
        b = oracle_mod.perturb(a, sig, xr, xp)
        assert abs(np.linalg.norm(b) - 1) < 1e-14
        th = math.acos(min(1.0, b @ a))
        th2.append(th * th / (2 * sig * sig))
        cu.append((b @ u) / sig)
    assert sst.kstest(th2, "expon").pvalue > 1e-3
    assert sst.kstest(cu, "norm").pvalue > 1e-3


def test_gaussian_sun_spread(oracle_mod):
    w = scenes.parabolic_dish()
    s = oracle_mod.Scene(w.elements)
    s.set_sun(w.sun_vector, w.sun_param, SUN_GAUSSIAN)
    assert s.sun_plane()["box_half_angle"] == pytest.approx(4 * w.sun_param)
    wv = -w.sun_vector / np.linalg.norm(w.sun_vector)
    th2 = []
    for i in range(3000):
        _, d = s.sample_sun(i)
        th = math.acos(min(1.0, d @ wv))
        th2.append(th * th / (2 * w.sun_param ** 2))
    assert sst.kstest(th2, "expon").pvalue > 1e-3


def _receiver_only_scene(tilt_mirror=False):
    """A 2 x 2 m flat receiver at the origin facing the zenith sun, and a small
    mirror high up and far to the side so the sun plane sits above the receiver."""
    e = scenes.make_elements(2)
    e[0]["origin"], e[0]["aim_point"] = [60.0, 0.0, 30.0], [60.0, 0.0, 40.0]
    e[0]["dim_x"], e[0]["dim_y"], e[0]["surface"], e[0]["role"] = 1.0, 1.0, SURFACE_FLAT, ROLE_MIRROR
    e[1]["origin"], e[1]["aim_point"] = [0.0, 0.0, 0.0], [0.0, 0.0, 5.0]
    e[1]["dim_x"], e[1]["dim_y"], e[1]["surface"], e[1]["role"] = 2.0, 2.0, SURFACE_FLAT, ROLE_RECEIVER
    return e


def test_uniform_illumination_closed_form(oracle_mod):
    """Point sun at the zenith on a sun-facing receiver: hits ~ Binomial(N, A_rec/A_box),
    bins uniform (chi-square), mean flux = DNI."""
    s = oracle_mod.Scene(_receiver_only_scene())
    s.set_sun([0, 0, 1], 0.0, dni=1000.0)
    pl = s.sun_plane()
    N = 400_000
    r = s.trace(N, max_depth=5, flux_nx=8, flux_ny=8, include_direct_sun=True, threads=8)
    p = 4.0 / pl["area"]
    hits = r["stats"][3]
    assert abs(hits - N * p) < 5 * math.sqrt(N * p * (1 - p))
    assert r["flux"].sum() == hits
    chi = sst.chisquare(r["flux"].ravel().astype(float))
    assert chi.pvalue > 1e-4
    power_per_ray = 1000.0 * pl["area"] / N
    mean_flux = r["flux"].mean() * power_per_ray / (2.0 / 8 * 2.0 / 8)
    assert mean_flux == pytest.approx(1000.0, rel=5 * math.sqrt((1 - p) / (N * p)))
    # SolTrace semantics (default): direct-sun hits are not binned
    r2 = s.trace(N, max_depth=5, flux_nx=8, flux_ny=8, threads=8)
    assert r2["flux"].sum() == 0 and r2["stats"][4] == r2["stats"][3]


def test_flat_mirror_images_onto_receiver(oracle_mod):
    """Point sun at the zenith, a flat mirror tilted so its reflection points
    at a large receiver: every mirror hit reaches the receiver (no truncation,
    reflectivity 1), and the landing spot is the mirror's image."""
    e = scenes.make_elements(2)
    # mirror at origin, normal bisecting zenith and +x: reflected rays travel along +x
    n = np.array([1.0, 0.0, 1.0]) / math.sqrt(2)
    e[0]["origin"], e[0]["aim_point"] = [0, 0, 0], n * 10
    e[0]["dim_x"], e[0]["dim_y"], e[0]["surface"], e[0]["role"] = 1.0, 1.0, SURFACE_FLAT, ROLE_MIRROR
    e[1]["origin"], e[1]["aim_point"] = [20, 0, 0], [10, 0, 0]        # faces -x
    e[1]["dim_x"], e[1]["dim_y"], e[1]["surface"], e[1]["role"] = 4.0, 4.0, SURFACE_FLAT, ROLE_RECEIVER
    s = oracle_mod.Scene(e)
    s.set_sun([0, 0, 1], 0.0)
    N = 100_000
    r = s.trace(N, max_depth=5, records=True, flux_nx=16, flux_ny=16, threads=4)
    st = r["stats"]
    assert st[2] > 0 and st[2] == st[3] - st[4] and st[10] == st[2]
    rec, ids = r["records"], r["record_elements"]
    via_mirror = ids[:, 1] == 0
    land = rec[via_mirror, 2, 1:]
    mir = rec[via_mirror, 1, 1:]
    assert np.allclose(land[:, 0], 20.0, atol=1e-9)
    assert np.allclose(land[:, 1], mir[:, 1], atol=1e-9)           # parallel rays along +x keep y, z
    assert np.allclose(land[:, 2], mir[:, 2], atol=1e-9)


def test_reflectivity_absorbs_expected_fraction(oracle_mod):
    w = scenes.heliostat_field(1000, 200_000)
    s = oracle_mod.Scene(w.elements)
    s.set_sun(w.sun_vector, w.sun_param)
    r = s.trace(w.num_rays, max_depth=5, threads=8)
    m, a = r["stats"][2], r["stats"][5]
    p = 1 - 0.93
    assert abs(a - m * p) < 5 * math.sqrt(m * p * (1 - p))


def test_conservation_and_thread_independence(oracle_mod):
    for w in (scenes.three_heliostats(50_000), scenes.parabolic_dish(50_000), scenes.heliostat_field(2000, 50_000)):
        s = oracle_mod.Scene(w.elements)
        s.set_sun(w.sun_vector, w.sun_param, w.sun_shape)
        a = s.trace(w.num_rays, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny, threads=1)
        b = s.trace(w.num_rays, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny, threads=7)
        for k in ("flux", "element_hits", "stats"):
            assert np.array_equal(a[k], b[k])
        st = a["stats"]
        assert ray_outcomes_balance(st)
        assert st[0] == w.num_rays
        assert a["flux"].sum() == st[10]
        assert a["element_hits"].sum() == st[2] + st[3]
        # chunking: two halves == one call
        h1 = s.trace(w.num_rays, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny,
                     ray_begin=0, ray_count=w.num_rays // 2, threads=3)
        h2 = s.trace(w.num_rays, max_depth=w.max_depth, flux_nx=w.flux_nx, flux_ny=w.flux_ny,
                     ray_begin=w.num_rays // 2, threads=3)
        assert np.array_equal(h1["flux"] + h2["flux"], a["flux"])
        assert np.array_equal(h1["stats"] + h2["stats"], a["stats"])


def test_depth_cap_semantics(oracle_mod):
    """MAX_TRACE_DEPTH (cuda/Soltrace.h:11): interaction k is recorded in slot k
    only while k < max_depth (cuda/materials.cu:56, 114)."""
    w = scenes.three_heliostats(20_000)
    s = oracle_mod.Scene(w.elements)
    s.set_sun(w.sun_vector, w.sun_param)
    full = s.trace(w.num_rays, max_depth=5, records=True, threads=4)
    one = s.trace(w.num_rays, max_depth=2, records=True, threads=4)
    # with max_depth 2 only the first interaction is recorded; mirror-then-receiver paths truncate
    assert one["stats"][3] == full["stats"][4]              # only direct receiver hits remain
    ids = full["record_elements"]
    second = np.sum((ids[:, 1] < 3) & (ids[:, 2] != 0xFFFFFFFF))
    assert one["stats"][8] == second                        # mirror paths that hit something next are cut
    assert np.all(one["record_elements"][:, 0] == 0xFFFFFFFE)
    assert np.array_equal(one["records"][:, :2], full["records"][:, :2])


def test_records_layout(oracle_mod):
    w = scenes.three_heliostats(5_000)
    s = oracle_mod.Scene(w.elements)
    s.set_sun(w.sun_vector, w.sun_param)
    r = s.trace(w.num_rays, max_depth=5, records=True, threads=2)
    rec, ids = r["records"], r["record_elements"]
    assert np.all(rec[:, 0, 0] == 0.0) and np.all(ids[:, 0] == 0xFFFFFFFE)
    empty = ids == 0xFFFFFFFF
    assert np.all(rec[empty] == 0.0)
    # tags: 1 for mirrors, 2 for the receiver (element 3)
    assert np.all(rec[ids == 3][:, 0] == 2.0)
    assert np.all(rec[(ids < 3)][:, 0] == 1.0)


def test_stratified_cells_cover_the_grid_once_per_sweep(oracle_mod):
    """Position sampling mode 2: the 65,536 groups of one sweep visit every cell of the
    256 x 256 grid exactly once (a bijection), in an order that depends on seed, job, sweep."""
    cells0 = [oracle_mod.stratified_cell(g * 64, seed=3, job=1) for g in range(65536)]
    assert sorted(cells0) == list(range(65536))
    cells1 = [oracle_mod.stratified_cell((65536 + g) * 64, seed=3, job=1) for g in range(65536)]
    assert sorted(cells1) == list(range(65536)) and cells0 != cells1
    assert cells0 != [oracle_mod.stratified_cell(g * 64, seed=4, job=1) for g in range(65536)]
    # 64 consecutive rays share their group's cell
    assert len({oracle_mod.stratified_cell(128 + k, 3, 1) for k in range(64)}) == 1


def test_stratified_positions_uniform(oracle_mod):
    from optixsoltracedemos_amd._abi import POS_STRATIFIED
    w = scenes.three_heliostats()
    s = oracle_mod.Scene(w.elements)
    s.set_sun(w.sun_vector, w.sun_param, position_sampling=POS_STRATIFIED)
    pl = s.sun_plane()
    e1, e2 = pl["v1"] - pl["v0"], pl["v3"] - pl["v0"]
    ab = []
    for i in range(0, 64 * 4000, 61):          # spread over many groups
        p, _ = s.sample_sun(i, seed=9)
        r = p - pl["v0"]
        ab.append((r @ e1 / (e1 @ e1), r @ e2 / (e2 @ e2)))
    ab = np.array(ab)
    for k in range(2):
        assert sst.kstest(ab[:, k], "uniform").pvalue > 1e-3
    # same flux statistics as i.i.d. sampling within Monte Carlo noise
    a = s.trace(400_000, max_depth=5, flux_nx=8, flux_ny=8, threads=8)
    s.set_sun(w.sun_vector, w.sun_param)
    b = s.trace(400_000, max_depth=5, flux_nx=8, flux_ny=8, threads=8)
    ha, hb = int(a["stats"][3]), int(b["stats"][3])
    assert abs(ha - hb) < 5 * math.sqrt(ha + hb)


def test_halton_range_past_2_32_is_rejected(oracle_mod):
    """Halton positions take the 32-bit launch index of cuda/sun.cu:94-134: a ray
    range that passes 2^32 is an error, one that ends at 2^32 is traced."""
    w = scenes.three_heliostats()
    s = oracle_mod.Scene(w.elements)
    s.set_sun(w.sun_vector, w.sun_param, position_sampling=POS_HALTON)
    with pytest.raises(RuntimeError):
        s.trace((1 << 32) + 64, ray_begin=(1 << 32) - 64, ray_count=128, threads=1)
    r = s.trace(1 << 32, ray_begin=(1 << 32) - 128, ray_count=128, threads=1)
    assert r["stats"][0] == 128
    s.set_sun(w.sun_vector, w.sun_param)   # Philox positions take the 64-bit index
    assert s.trace((1 << 32) + 64, ray_begin=(1 << 32) - 64, ray_count=128, threads=1)["stats"][0] == 128
