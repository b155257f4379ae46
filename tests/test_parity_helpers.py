"""The comparison helpers the GPU parity tests rely on, checked on planted
differences (CPU)."""
import numpy as np

from tests.parity import compare_directions, compare_records, flux_l1

EMPTY, ORIGIN = 0xFFFFFFFF, 0xFFFFFFFE


def _rays():
    rec = np.zeros((4, 4, 4))
    ids = np.full((4, 4), EMPTY, np.uint32)
    rec[:, 0, 1:] = [0.0, 0.0, 1000.0]   # sun-plane origin
    rec[:, 1, 1:] = [0.0, 0.0, 0.0]      # mirror hit
    rec[:, 2, 1:] = [10.0, 0.0, 5.0]     # receiver hit
    rec[:, 1:3, 0] = [1.0, 2.0]
    ids[:, 0], ids[:, 1], ids[:, 2] = ORIGIN, 3, 7
    return rec, ids


def test_directions_zero_on_identical_records():
    rec, ids = _rays()
    assert np.all(compare_directions(rec, ids, rec.copy(), ids) == 0.0)


def test_directions_measure_a_tilted_segment():
    rec, ids = _rays()
    o = rec.copy()
    o[2, 2, 1:] = [10.0, 1e-3 * np.hypot(10.0, 5.0), 5.0]   # second segment turned by ~1e-3 rad
    ang = compare_directions(rec, ids, o, ids)
    assert ang[0] == ang[1] == ang[3] == 0.0
    assert abs(ang[2] - 1e-3) < 1e-6


def test_directions_skip_rays_whose_elements_differ_and_empty_slots():
    rec, ids = _rays()
    o = rec.copy()
    oid = ids.copy()
    oid[1, 2] = 8                          # another element: ray 1 not compared
    o[1, 2, 1:] = [0.0, 10.0, 5.0]
    ang = compare_directions(rec, ids, o, oid)
    assert len(ang) == 3 and np.all(ang == 0.0)


def test_records_and_flux_helpers_see_planted_errors():
    rec, ids = _rays()
    o = rec.copy()
    o[0, 1, 1] += 0.25
    frac, err, same = compare_records(rec, ids, o, ids)
    assert frac == 0.0 and abs(err - 0.25) < 1e-12 and same.all()
    a = np.zeros((4, 4), np.int64)
    a[1, 1] = 100
    b = a.copy()
    b[1, 1], b[1, 2] = 99, 1
    assert abs(flux_l1(a, b) - 0.02) < 1e-12
