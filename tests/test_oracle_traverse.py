"""Pins for the CPU oracle's traversal (no GPU): the closest hit equals a
brute-force loop over every triangle with the same Moller-Trumbore test
(Core/Triangle.h:7-52), and any-hit visibility equals "no triangle hit
closer than the limit"."""
import numpy as np
import pytest

import oracle
from cavalry_amd import scenes


def brute_force(tris, rays):
    """All (ray, triangle) hits, float32 arithmetic in the reference's order."""
    f = np.float32
    v0 = tris[:, 0:3]
    e1 = tris[:, 3:6] - v0
    e2 = tris[:, 6:9] - v0
    o = rays[:, None, 0:3]
    d = rays[:, None, 3:6]

    def cross(a, b):
        return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                         a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                         a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)

    def dot(a, b):
        return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]

    with np.errstate(all="ignore"):
        tvec = o - v0[None]
        pvec = cross(np.broadcast_to(d, tvec.shape), e2[None])
        det = dot(e1[None], pvec)
        inv = f(1) / det
        u = dot(tvec, pvec) * inv
        qvec = cross(tvec, e1[None])
        v = dot(np.broadcast_to(d, tvec.shape), qvec) * inv
        t = dot(e2[None], qvec) * inv
        hit = (det != 0) & (u >= 0) & (u <= 1) & (v >= 0) & (u + v <= 1) & (t > 0)
    return hit, t


@pytest.mark.parametrize("rounds", [0, 2])
def test_closest_hit_equals_brute_force(rounds):
    tris = scenes.gmm_triangles(1500, seed=21, edge=0.03)
    rays = scenes.random_rays(3000, seed=22, lo=0.1, hi=0.9)
    _, _, nd = oracle.lbvh_build(tris, 30)
    if rounds:
        oracle.treelet_optimize(nd, rounds)
    got = oracle.trace_closest(nd, tris, rays)
    hit, t = brute_force(tris, rays)
    tt = np.where(hit, t, np.inf)
    best = tt.min(axis=1)
    has = np.isfinite(best)
    assert has.sum() > 100
    np.testing.assert_array_equal(got["prim"] >= 0, has)
    np.testing.assert_array_equal(got["t"][has], best[has])
    # the reported triangle is one of those at the minimum distance
    rows = np.nonzero(has)[0]
    assert (tt[rows, got["prim"][rows]] == best[rows]).all()


def test_any_hit_equals_brute_force():
    tris = scenes.gmm_triangles(1500, seed=31, edge=0.03)
    rays = scenes.random_rays(3000, seed=32, lo=0.1, hi=0.9)
    lim = scenes.shadow_limits(3000, seed=33) * np.float32(0.3)
    lim[::5] = np.inf
    _, _, nd = oracle.lbvh_build(tris, 30)
    vis = oracle.trace_occluded(nd, tris, rays, lim)
    hit, t = brute_force(tris, rays)
    blocked = (hit & (t < lim[:, None])).any(axis=1)
    np.testing.assert_array_equal(vis == 0, blocked)
    assert blocked.sum() > 50 and (~blocked).sum() > 50


def test_mesh_exclusion():
    tris = scenes.gmm_triangles(800, seed=41, edge=0.03)
    rays = scenes.random_rays(2000, seed=42, lo=0.1, hi=0.9)
    lim = np.full(2000, np.inf, np.float32)
    mesh = (np.arange(800) % 5).astype(np.int32)
    src = np.full(2000, 1, np.int32)
    dst = np.full(2000, 3, np.int32)
    _, _, nd = oracle.lbvh_build(tris, 30)
    vis = oracle.trace_occluded(nd, tris, rays, lim, mesh, src, dst)
    hit, _ = brute_force(tris, rays)
    keep = (mesh != 1) & (mesh != 3)
    np.testing.assert_array_equal(vis == 0, (hit & keep[None]).any(axis=1))
