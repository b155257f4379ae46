"""The conservative tests that build candidate lists, checked on the CPU
against brute-force ray sampling (tests/cpp/cull_check.cpp over st_common.h,
the same functions the list-building kernels run):

* cone_meets_box (the BVH walk of the reflection lists: the reflection cone
  from a mirror's box against candidate boxes): no box around a point such a
  ray passes through is dropped, and every dropped box is missed by sampled rays;
* rect_sweep_misses (first- and second-reflection lists of flat mirrors): no
  box around a point that a ray from the mirror's rectangle, within the
  reflection cone, passes through is ever dropped, and every dropped box is
  missed by thousands of such rays; rect_sweep_misses_rect, the same for flat
  candidates' rectangles;
* grid_range / cell_meets (sun-plane grid cells): every ray that hits a flat
  primitive's rectangle, or enters a curved primitive's traversal box, traced
  back within the sun cone to the sun plane, starts in a listed cell;
* sphere_cone_reaches (re-aim candidates of batched sweeps, k_reaim_lists):
  spheres around points that a bisector-aimed mirror's reflections pass
  through are never dropped, with the cone's axis taken towards the target as
  the kernel takes it, and dropped spheres are missed by sampled reflections.

Each check also runs with a planted mistake (a cone narrower than the real one
when the test is built) and must then report violations, so the sampling has
the power to catch an unsound test.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cull") / "cull_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "optixsoltracedemos_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "cull_check.cpp"), "-o", out], check=True)
    return out


def run(exe, mode, seed, scale):
    line = subprocess.run([exe, mode, str(seed), str(scale)], stdout=subprocess.PIPE, text=True,
                          check=True).stdout.split()
    assert line[0] == mode
    return {line[i]: int(line[i + 1]) for i in range(1, len(line), 2)}


@pytest.mark.parametrize("seed", [1, 2])
def test_reflection_cone_walk_is_conservative(exe, seed):
    r = run(exe, "cone", seed, 1.0)
    assert r["reachable_boxes"] > 100_000 and r["reachable_dropped"] == 0, r
    assert r["dropped"] > 10_000 and r["dropped_but_hit"] == 0, r


def test_reflection_cone_check_catches_a_narrow_cone(exe):
    r = run(exe, "cone", 1, 0.5)
    assert r["reachable_dropped"] > 0 and r["dropped_but_hit"] > 0, r


@pytest.mark.parametrize("seed", [1, 2])
def test_reflection_sweep_culling_is_conservative(exe, seed):
    r = run(exe, "sweep", seed, 1.0)
    assert r["reachable_boxes"] > 100_000 and r["reachable_dropped"] == 0, r
    assert r["dropped"] > 10_000 and r["dropped_but_hit"] == 0, r


def test_reflection_sweep_check_catches_a_narrow_cone(exe):
    r = run(exe, "sweep", 1, 0.5)
    assert r["reachable_dropped"] > 0 and r["dropped_but_hit"] > 0, r


@pytest.mark.parametrize("seed", [1, 2])
def test_reflection_sweep_rectangle_culling_is_conservative(exe, seed):
    """rect_sweep_misses_rect (flat candidates culled by their rectangle, not
    their box): rectangles through points that rays pass are never dropped, and
    dropped rectangles are missed by every sampled ray; some rectangles whose
    box the box test keeps are dropped."""
    r = run(exe, "sweeprect", seed, 1.0)
    assert r["reachable"] > 100_000 and r["reachable_dropped"] == 0, r
    assert r["dropped"] > 10_000 and r["dropped_but_hit"] == 0 and r["dropped_box_kept"] > 0, r


def test_reflection_sweep_rectangle_check_catches_a_narrow_cone(exe):
    r = run(exe, "sweeprect", 1, 0.5)
    assert r["reachable_dropped"] > 0 and r["dropped_but_hit"] > 0, r


@pytest.mark.parametrize("seed", [1, 2])
def test_grid_cells_of_flat_primitives_are_conservative(exe, seed):
    r = run(exe, "grid", seed, 1.0)
    assert r["hits"] > 500_000 and r["unlisted"] == 0, r
    assert r["listed_cells"] < r["range_cells"], r   # the parallelogram test drops cells


def test_grid_check_catches_a_narrow_cone(exe):
    r = run(exe, "grid", 1, 0.3)
    assert r["unlisted"] > 0, r


@pytest.mark.parametrize("seed", [1, 2])
def test_grid_cells_of_curved_primitives_are_conservative(exe, seed):
    r = run(exe, "gridbox", seed, 1.0)
    assert r["hits"] > 500_000 and r["unlisted"] == 0, r


def test_grid_box_check_catches_a_narrow_cone(exe):
    r = run(exe, "gridbox", 1, 0.3)
    assert r["unlisted"] > 0, r


@pytest.mark.parametrize("seed", [1, 2])
def test_reaim_sphere_test_is_conservative(exe, seed):
    r = run(exe, "reaim", seed, 1.0)
    assert r["reachable"] > 100_000 and r["reachable_dropped"] == 0, r
    assert r["dropped"] > 10_000 and r["dropped_but_hit"] == 0, r


def test_reaim_check_catches_a_narrow_cone(exe):
    r = run(exe, "reaim", 1, 0.5)
    assert r["reachable_dropped"] > 0 and r["dropped_but_hit"] > 0, r
