"""Comparison helpers: HIP path (FP32) against the FP64 oracle on shared Philox draws."""
from __future__ import annotations

import numpy as np


def compare_records(g_rec, g_id, o_rec, o_id, per_ray=False):
    """Per-ray parity of hit records.

    Returns (id_mismatch_fraction, max position error over rays whose element
    sequences agree, mask of agreeing rays); with per_ray=True the second item
    is the array of per-ray maximum errors instead. Slot ids: 0xFFFFFFFE sun
    origin, 0xFFFFFFFF empty.
    """
    g_id = np.asarray(g_id).astype(np.uint32)
    o_id = np.asarray(o_id).astype(np.uint32)
    same = np.all(g_id == o_id, axis=1)
    frac = 1.0 - same.mean() if len(same) else 0.0
    g = np.asarray(g_rec, dtype=np.float64)[same]
    o = np.asarray(o_rec, dtype=np.float64)[same]
    if len(g) == 0:
        return frac, (np.zeros(0) if per_ray else 0.0), same
    tag_ok = np.all(g[..., 0] == o[..., 0])
    assert tag_ok, "record tags differ on rays whose element sequences agree"
    err = np.linalg.norm(g[..., 1:] - o[..., 1:], axis=-1)
    if per_ray:
        return frac, err.max(axis=1), same
    return frac, float(err.max()), same


def compare_directions(g_rec, g_id, o_rec, o_id):
    """Per-ray largest angle (radians) between the GPU's and the oracle's
    segment directions, over rays whose element sequences agree. A segment runs
    between consecutive recorded points: the sun-plane origin (slot 0), then
    every hit; its direction is the unit vector from one point to the next."""
    g_id = np.asarray(g_id).astype(np.uint32)
    o_id = np.asarray(o_id).astype(np.uint32)
    same = np.all(g_id == o_id, axis=1)
    ids = g_id[same]
    g = np.asarray(g_rec, dtype=np.float64)[same][..., 1:]
    o = np.asarray(o_rec, dtype=np.float64)[same][..., 1:]
    if len(g) == 0:
        return np.zeros(0)
    valid = ids != 0xFFFFFFFF
    seg = valid[:, :-1] & valid[:, 1:]

    def unit(v):
        n = np.linalg.norm(v, axis=-1, keepdims=True)
        return v / np.where(n > 0, n, 1.0)

    dg = unit(g[:, 1:] - g[:, :-1])
    do = unit(o[:, 1:] - o[:, :-1])
    cross = np.linalg.norm(np.cross(dg, do), axis=-1)
    ang = np.arctan2(cross, np.sum(dg * do, axis=-1))
    ang = np.where(seg, ang, 0.0)
    return ang.max(axis=1) if ang.shape[1] else np.zeros(len(ang))


def flux_l1(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    tot = max(a.sum(), b.sum(), 1.0)
    return float(np.abs(a - b).sum() / tot)


def ray_outcomes_balance(stats) -> bool:
    """Every ray ends exactly once: escaped, absorbed, back face, truncated or
    as a recorded receiver hit."""
    s = np.asarray(stats).astype(np.int64)
    return s[0] == s[6] + s[5] + s[7] + s[8] + s[3]
