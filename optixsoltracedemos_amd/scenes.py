"""Synthetic scenes for the five benchmark / parity workloads.

Element arrays use ``_abi.ELEMENT_DTYPE`` (the st_element_f64 layout). The
element conventions follow the reference: the element frame comes from
origin, aim point and zrot (lib/element.cpp:77-100), the receiver is the last
element (lib/geometry_manager.cpp:78-90), a cylinder's radius and length come
from its rectangular aperture (lib/element.cpp:150-160).

Sun vectors: x = east, y = north, z = up, azimuth clockwise from north,
``s = (sin az cos el, cos az cos el, sin el)``; the vector points to the sun.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._abi import (ELEMENT_DTYPE, ROLE_MIRROR, ROLE_RECEIVER, SURFACE_CYLINDER, SURFACE_FLAT,
                   SURFACE_PARABOLIC, SUN_GAUSSIAN, SUN_PILLBOX, POS_PHILOX)


def sun_vector(elevation_deg: float, azimuth_deg: float) -> np.ndarray:
    el, az = math.radians(elevation_deg), math.radians(azimuth_deg)
    return np.array([math.sin(az) * math.cos(el), math.cos(az) * math.cos(el), math.sin(el)])


def bisector_aim(origin: np.ndarray, target: np.ndarray, sun: np.ndarray, dist: float = 100.0) -> np.ndarray:
    """Aim point on the bisector of the sun direction and the direction to the
    target, so the mirror normal reflects the sun centre onto the target. The
    toy STINPUT's aim points satisfy this (data/stinput/toy_problem_parabolic.stinput:15-17)."""
    s = sun / np.linalg.norm(sun)
    t = (target - origin) / np.linalg.norm(target - origin)
    n = s + t
    n /= np.linalg.norm(n)
    return origin + dist * n


def make_elements(n: int) -> np.ndarray:
    e = np.zeros(n, dtype=ELEMENT_DTYPE)
    e["reflectivity"] = 1.0
    return e


@dataclass
class Workload:
    name: str
    elements: np.ndarray
    sun_vector: np.ndarray
    sun_shape: int
    sun_param: float
    num_rays: int
    max_depth: int
    flux_nx: int
    flux_ny: int
    seed: int = 0
    position_sampling: int = POS_PHILOX
    dni: float = 1000.0
    extra: dict = field(default_factory=dict)

    @property
    def receiver(self) -> int:
        return int(np.nonzero(self.elements["role"] == ROLE_RECEIVER)[0][-1])


# ---------------------------------------------------------------------------
def three_heliostats(num_rays: int = 1_000_000) -> Workload:
    """Three flat 1.0 x 1.95 m heliostats at (-5,5,0), (0,5,0), (5,5,0), bisector-aimed
    at a 2 x 2 m flat receiver at (0,0,10) facing the field (aim (0,5,0) as
    demos/demo_three_heliostats.cpp:133-134). Pillbox 4.65 mrad, el 60, az 180."""
    s = sun_vector(60.0, 180.0)
    rec_c = np.array([0.0, 0.0, 10.0])
    e = make_elements(4)
    for i, x in enumerate((-5.0, 0.0, 5.0)):
        o = np.array([x, 5.0, 0.0])
        e[i]["origin"] = o
        e[i]["aim_point"] = bisector_aim(o, rec_c, s)
        e[i]["dim_x"], e[i]["dim_y"] = 1.0, 1.95
        e[i]["surface"] = SURFACE_FLAT
        e[i]["role"] = ROLE_MIRROR
    e[3]["origin"] = rec_c
    e[3]["aim_point"] = [0.0, 5.0, 0.0]
    e[3]["dim_x"], e[3]["dim_y"] = 2.0, 2.0
    e[3]["surface"] = SURFACE_FLAT
    e[3]["role"] = ROLE_RECEIVER
    return Workload("three_heliostats_flat_receiver", e, s, SUN_PILLBOX, 4.65e-3, num_rays, 5, 64, 64)


def parabolic_dish(num_rays: int = 10_000_000) -> Workload:
    """One 5 x 5 m parabolic dish, focal length 3 m (c = 1/(2f)), axis along the
    sun; a capped cylinder r 0.25 m, length 0.5 m centred on the focus with its
    axis across the dish axis. Gaussian sun sigma 2.73 mrad, slope 2 mrad,
    specular 0.5 mrad, at most 2 interactions (max_depth 3)."""
    s = sun_vector(60.0, 180.0)
    f = 3.0
    c = 1.0 / (2.0 * f)
    e = make_elements(2)
    e[0]["origin"] = [0.0, 0.0, 0.0]
    e[0]["aim_point"] = 100.0 * s
    e[0]["dim_x"], e[0]["dim_y"] = 5.0, 5.0
    e[0]["curv_x"], e[0]["curv_y"] = c, c
    e[0]["surface"] = SURFACE_PARABOLIC
    e[0]["role"] = ROLE_MIRROR
    e[0]["slope_error"] = 2.0e-3
    e[0]["spec_error"] = 0.5e-3
    focus = f * s
    e[1]["origin"] = focus
    e[1]["aim_point"] = focus - s   # local z = -s, so the axis (local y) is across the dish axis
    e[1]["dim_x"], e[1]["dim_y"] = 0.5, 0.5
    e[1]["surface"] = SURFACE_CYLINDER
    e[1]["role"] = ROLE_RECEIVER
    return Workload("parabolic_dish_cylinder_receiver", e, s, SUN_GAUSSIAN, 2.73e-3, num_rays, 3, 128, 64)


def _stagger_rings(d: float, r_min: float, r_max: float):
    """Zoned radial stagger with centre spacing >= d.

    Inside a zone every ring holds the same count m and odd rings are offset by
    half a pitch; the radial step from a ring of pitch p is sqrt(d^2 - (p/2)^2),
    so staggered neighbours on adjacent rings stay d apart. When the pitch grows
    past 1.5 d a new zone starts one full spacing d further out, with the count
    re-derived from the new radius (1.5 packs 10,000 centres into 100-800 m at
    d = 14.38 m)."""
    rings = []
    r = r_min
    m = int(math.floor(2 * math.pi * r / d))
    k = 0
    while r <= r_max + 1e-9:
        p = 2 * math.pi * r / m
        rings.append((r, m, k % 2))
        if p < 1.5 * d:
            r += math.sqrt(d * d - 0.25 * p * p)
            k += 1
        else:
            r += d
            m = int(math.floor(2 * math.pi * r / d))
            k = 0
    return rings


def radial_stagger_positions(n: int, r_min: float, r_max: float, seed: int, jitter: float = 0.05) -> np.ndarray:
    """n heliostat centres on zoned staggered rings between r_min and r_max.

    The spacing d is the largest for which the rings hold n positions (found by
    bisection); surplus positions are dropped from the outermost ring and each
    centre is jittered by up to `jitter` m in x and y from a Philox stream of
    `seed`. The caller checks the minimum spacing."""
    lo, hi = 1.0, 2.0 * math.sqrt(math.pi * (r_max ** 2 - r_min ** 2) / n)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if sum(m for _, m, _ in _stagger_rings(mid, r_min, r_max)) >= n:
            lo = mid
        else:
            hi = mid
    rings = _stagger_rings(lo, r_min, r_max)
    pts = []
    left = n
    for r, m, odd in rings:
        if left <= 0:
            break
        step = 2 * math.pi / m
        th = (0.5 * step if odd else 0.0) + step * np.arange(min(m, left))
        pts.append(np.stack([r * np.sin(th), r * np.cos(th)], axis=1))
        left -= len(th)
    xy = np.concatenate(pts, axis=0)
    rng = np.random.Generator(np.random.Philox(seed))
    return xy + rng.uniform(-jitter, jitter, size=xy.shape)


def min_spacing(xy: np.ndarray) -> float:
    from scipy.spatial import cKDTree
    d, _ = cKDTree(xy).query(xy, k=2)
    return float(d[:, 1].min())


def heliostat_field(n_heliostats: int = 10_000, num_rays: int = 100_000_000, seed: int = 0,
                    elevation: float = 60.0, azimuth: float = 180.0, r_min: float = 100.0,
                    r_max: float | None = None) -> Workload:
    """Radial-stagger field of flat 10 x 10 m heliostats (z = 0) around a 150 m
    tower carrying a vertical capped-cylinder receiver r 8 m, height 20 m.
    Bisector aim at the receiver centre, pillbox 4.65 mrad, slope 1.5 mrad,
    reflectivity 0.93. 10,000 heliostats use rings from 100 to 800 m."""
    if r_max is None:
        r_max = 800.0 if n_heliostats >= 10_000 else max(r_min + 30.0, math.sqrt(r_min ** 2 + n_heliostats * 240.0 / math.pi))
    s = sun_vector(elevation, azimuth)
    rec_c = np.array([0.0, 0.0, 150.0])
    xy = radial_stagger_positions(n_heliostats, r_min, r_max, seed)
    e = make_elements(n_heliostats + 1)
    org = np.zeros((n_heliostats, 3))
    org[:, :2] = xy
    t = rec_c[None, :] - org
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    nrm = s[None, :] + t
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    m = e[:n_heliostats]
    m["origin"] = org
    m["aim_point"] = org + 100.0 * nrm
    m["dim_x"], m["dim_y"] = 10.0, 10.0
    m["surface"] = SURFACE_FLAT
    m["role"] = ROLE_MIRROR
    m["slope_error"] = 1.5e-3
    m["reflectivity"] = 0.93
    r = e[n_heliostats]
    r["origin"] = rec_c
    r["aim_point"] = rec_c + np.array([0.0, -1.0, 0.0])  # frame x=(1,0,0), z=(0,-1,0): axis = +z
    r["dim_x"], r["dim_y"] = 16.0, 20.0
    r["surface"] = SURFACE_CYLINDER
    r["role"] = ROLE_RECEIVER
    return Workload(f"field_{n_heliostats}_heliostats_cylinder_receiver", e, s, SUN_PILLBOX, 4.65e-3,
                    num_rays, 5, 256, 128, seed=seed, extra={"target": rec_c})


def annual_sweep(steps: int = 96, seed: int = 0):
    """96 suns, elevation U[5,85] deg, azimuth U[60,300] deg, budgets log-uniform
    in [1e6, 5e7] rays, all drawn from a Philox stream of `seed`."""
    rng = np.random.Generator(np.random.Philox(seed + 1))
    el = rng.uniform(5.0, 85.0, steps)
    az = rng.uniform(60.0, 300.0, steps)
    budgets = np.exp(rng.uniform(math.log(1e6), math.log(5e7), steps)).astype(np.uint64)
    suns = np.stack([sun_vector(a, b) for a, b in zip(el, az)])
    return suns, budgets


WORKLOADS = {
    "three_heliostats": three_heliostats,
    "parabolic_dish": parabolic_dish,
    "field_10k": heliostat_field,
}
