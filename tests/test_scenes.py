"""Synthetic workload builders."""
import math

import numpy as np

from optixsoltracedemos_amd import scenes
from optixsoltracedemos_amd._abi import ROLE_RECEIVER


def test_field_10k_layout():
    w = scenes.heliostat_field(10_000, 1000, seed=0)
    m = w.elements[:-1]
    assert len(m) == 10_000
    xy = m["origin"][:, :2]
    r = np.hypot(xy[:, 0], xy[:, 1])
    assert r.min() > 100 - 0.1 and r.max() < 800 + 0.1
    assert scenes.min_spacing(xy) >= 14.2             # mirror diagonal 10*sqrt(2)
    assert w.elements[-1]["role"] == ROLE_RECEIVER and w.receiver == 10_000
    assert np.allclose(w.elements[-1]["origin"], [0, 0, 150])


def test_field_seeded():
    a = scenes.heliostat_field(1000, 10, seed=1).elements["origin"]
    b = scenes.heliostat_field(1000, 10, seed=1).elements["origin"]
    c = scenes.heliostat_field(1000, 10, seed=2).elements["origin"]
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    for n in (100, 1000):
        xy = scenes.heliostat_field(n, 10).elements["origin"][:-1, :2]
        assert scenes.min_spacing(xy) >= 14.2


def test_bisector_aims():
    w = scenes.heliostat_field(1000, 10)
    s = w.sun_vector / np.linalg.norm(w.sun_vector)
    tgt = np.array([0, 0, 150.0])
    for e in w.elements[:-1][:50]:
        n = (e["aim_point"] - e["origin"]) / np.linalg.norm(e["aim_point"] - e["origin"])
        t = (tgt - e["origin"]) / np.linalg.norm(tgt - e["origin"])
        r = 2 * (s @ n) * n - s                       # reflected sun centre
        assert np.allclose(r, t, atol=1e-12)


def test_sun_vector_convention():
    assert np.allclose(scenes.sun_vector(90, 0), [0, 0, 1], atol=1e-15)
    assert np.allclose(scenes.sun_vector(0, 90), [1, 0, 0], atol=1e-15)        # east
    assert np.allclose(scenes.sun_vector(60, 180), [0, -0.5, math.sqrt(3) / 2], atol=1e-15)


def test_annual_sweep_ranges():
    suns, budgets = scenes.annual_sweep(96, seed=0)
    assert suns.shape == (96, 3) and budgets.shape == (96,)
    el = np.degrees(np.arcsin(suns[:, 2]))
    assert el.min() >= 5 and el.max() <= 85
    assert budgets.min() >= 1e6 and budgets.max() <= 5e7
