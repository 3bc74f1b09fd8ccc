"""Pins for the oracle's sphere lights (no GPU).

A sphere shape that emits (ShapeMesh/Sphere.cu:84-88 sets shapeType and
shapeParams) is sampled within the cone it subtends from the shading point
(TriangleMesh::sampleSphere, Core/TriangleMesh.h:155-211), and a BSDF-sampled
ray that hits it is weighted with that cone's pdf (TriangleMesh::pdf,
Core/TriangleMesh.h:214-232). Closed form: a matte floor of albedo rho under
a sphere of radiance L and radius r fully above a floor point's horizon
receives the irradiance pi L r^2 cos(beta) / D^2 (D the distance to the
centre, beta its angle from the normal), so the floor's radiance is
rho L r^2 cos(beta) / D^2:

* directlighting draws only the cone, so its estimate has exactly that
  expectation;
* path at max depth 2 (a black emitter, so nothing else contributes) adds
  the cone sample and the BSDF sample with power-heuristic weights; the sum
  is unbiased only when the BSDF hit uses the cone pdf (the area pdf of the
  tessellated sphere gives weights that do not sum to 1: +10% here). The
  BSDF rays meet the 2048-triangle tessellation, not the sphere, which is
  within the tolerance;
* rlpath at max depth 2 combines the cone sample with the direction drawn
  from the learnt 128-bin Q distribution (power heuristic against the
  bin's pdf, ReinforcementLearningPathTracing.cu:416,644); the
  weights must again sum to 1 whatever the table has learnt, so the same
  closed form holds;
* a small far sphere (sin^2 of the cone's half-angle below sin^2 1.5 deg)
  goes through the reference's small-angle form and keeps the same answer.
"""
import numpy as np
import pytest

import oracle
from cavalry_amd import scenes
from cavalry_amd.camera import look_at


def cam_of(sc):
    return look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)


def _floor_pixels(sc, r, margin=1.0):
    """Closed-form floor radiance per pixel (NaN where a pixel's filter
    support might see the sphere or the sphere is not above the horizon)."""
    H, W = sc.height, sc.width
    cam = cam_of(sc)
    m = np.ctypeslib.as_array(cam.cam_to_world).reshape(4, 4).T.astype(np.float64)
    o = (m @ np.array([0, 0, 0, 1.0]))[:3]
    ys, xs = np.mgrid[0:H + 1, 0:W + 1].astype(np.float64) - 0.5
    ok = np.ones((H, W), bool)
    centre = np.array([0.0, 1.0, 0.0])
    for dy in (-margin, 0.0, margin):
        for dx in (-margin, 0.0, margin):
            p = np.stack([xs[:-1, :-1] + 0.5 + dx - W / 2, H / 2 - (ys[:-1, :-1] + 0.5 + dy),
                          np.full((H, W), cam.depth), np.ones((H, W))], -1) @ m.T
            d = p[..., :3] / p[..., 3:] - o
            d /= np.linalg.norm(d, axis=-1, keepdims=True)
            down = d[..., 1] < -1e-3
            t = -o[1] / np.where(down, d[..., 1], -1)
            hit = o + t[..., None] * d
            # a ray that meets the sphere (slightly enlarged) before the floor
            oc = o - centre
            bq = (d * oc).sum(-1)
            disc = bq ** 2 - ((oc * oc).sum() - (1.02 * r) ** 2)
            sphere_first = (disc > 0) & (-bq - np.sqrt(np.maximum(disc, 0)) < t)
            ok &= down & ~sphere_first & (np.abs(hit[..., 0]) < 3) & (np.abs(hit[..., 2]) < 3)
            if dx == 0 and dy == 0:
                P = hit
    D = np.linalg.norm(centre - P, axis=-1)
    cos_b = (centre[1] - P[..., 1]) / D
    ok &= np.arcsin(np.minimum(r / D, 1)) + np.arccos(np.clip(cos_b, -1, 1)) < np.pi / 2 - 0.02
    return np.where(ok, cos_b / D ** 2, np.nan)


@pytest.mark.parametrize("r", [0.7, 0.02])
@pytest.mark.parametrize("integrator", ["direct", "path", "rlpath"])
def test_sphere_light_floor_closed_form(integrator, r):
    rho, L = 0.5, 2.0 * (0.7 / r) ** 2
    sc = scenes.sphere_light_plane(width=64, height=48, albedo=rho, radiance=L, radius=r, max_depth=2)
    cam = cam_of(sc)
    if integrator == "direct":
        rgb, w, _ = oracle.render_direct(sc, 64, cam)
    elif integrator == "rlpath":
        rgb, w, _ = oracle.render_rlpath(sc, 64, cam, max_depth=2)
    else:
        rgb, w, _ = oracle.render_path(sc, 64, cam)
    img = (rgb / w[..., None])[..., 0]
    g = _floor_pixels(sc, r)
    sel = np.isfinite(g)
    assert sel.sum() > 300
    ratio = img[sel] / (rho * L * r * r * g[sel])
    assert ratio.mean() == pytest.approx(1.0, abs=0.01), ratio.mean()


def test_sphere_light_cone_against_triangle_sampler():
    """The same sphere marked as a plain mesh goes through the triangle
    sampler (a uniform triangle, a point in it, the area pdf): a different
    image of the same expectation in directlighting, with more variance."""
    rho, L, r = 0.5, 2.0, 0.7
    sc = scenes.sphere_light_plane(width=64, height=48, albedo=rho, radiance=L, radius=r, max_depth=1)
    plain = scenes.sphere_light_plane(width=64, height=48, albedo=rho, radiance=L, radius=r, max_depth=1)
    plain.meshes["sphere"] = 0
    cam = cam_of(sc)
    a, wa, _ = oracle.render_direct(sc, 64, cam)
    b, wb, _ = oracle.render_direct(plain, 64, cam)
    assert not np.array_equal(a, b)
    g = _floor_pixels(sc, r)
    sel = np.isfinite(g)
    ia, ib = (a / wa[..., None])[..., 0], (b / wb[..., None])[..., 0]
    # the cone sample has no variance from the light's shape here; the
    # triangle sampler does
    ra = ia[sel] / (rho * L * r * r * g[sel])
    rb = ib[sel] / (rho * L * r * r * g[sel])
    assert ra.std() < 0.5 * rb.std()
    assert rb.mean() == pytest.approx(1.0, abs=0.03)
