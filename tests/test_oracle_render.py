"""Pins for the CPU oracle's `path` integrator (no GPU).

* Halton sampler: the digit-permutation table and radical inverses equal an
  independent construction from the published Microsoft C runtime rand()
  recurrence (tests/golden/halton_fixture.json);
* furnace: inside a closed box whose every face is a diffuse area light of
  radiance Le and albedo rho, the radiance is Le / (1 - rho) everywhere; the
  estimator (NEE + MIS + Russian roulette) must agree to within its noise;
* the same furnace holds for rlpath (learnt light choice and directions);
* film: weights count exactly the samples whose box-filter support covers
  each pixel (4 per sample away from the edges).
"""
import json
import os

import numpy as np

import oracle
from cavalry_amd import scenes
from cavalry_amd.camera import look_at

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_halton_table_and_radical_inverse():
    with open(os.path.join(GOLDEN, "halton_fixture.json")) as f:
        fx = json.load(f)
    primes, start, perm = oracle.halton_table()
    assert primes[-1] == fx["prime_1023"] and int(primes.sum()) == fx["prime_sum"]
    for dim, p in enumerate(primes[:6]):
        head = fx["perm_head"][str(int(p))]
        np.testing.assert_array_equal(perm[start[dim]:start[dim] + p], head)
    for c in fx["radical_inverse"]:
        assert oracle.halton(c["index"], c["dim"]) == np.float32(c["value"]), c


def furnace_scene(rho=0.5, le=1.0, n=24):
    """Unit cube, six light faces with inward normals."""
    b = scenes._Builder()
    faces = [
        scenes.quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)),   # y=0, +y
        scenes.quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)),   # y=1, -y
        scenes.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)),   # z=0, +z
        scenes.quad((0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)),   # z=1, -z
        scenes.quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)),   # x=0, +x
        scenes.quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)),   # x=1, -x
    ]
    for f in faces:
        b.mesh(f, kd=(rho, rho, rho), emission=(le, le, le))
    tris, meshes = b.arrays()
    return scenes.SceneArrays(tris, meshes, (0.5, 0.5, 0.1), (0.5, 0.5, 0.9), (0, 1, 0), 60.0,
                              n, n, max_depth=24)


def test_face_normals_point_inward():
    sc = furnace_scene()
    t = sc.tris.reshape(-1, 3, 3)
    nrm = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    to_center = 0.5 - t.mean(axis=1)
    assert (np.einsum("ij,ij->i", nrm, to_center) > 0).all()


def test_furnace_radiance_is_le_over_one_minus_rho():
    for rho in (0.3, 0.6):
        sc = furnace_scene(rho=rho)
        cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
        rgb, w, _ = oracle.render_path(sc, 32, cam)
        img = rgb / w[..., None]
        expect = 1.0 / (1.0 - rho)
        mean = float(img.mean())
        assert abs(mean - expect) < 0.02 * expect, (rho, mean, expect)


def test_film_weights_count_box_filter_support():
    sc = scenes.cornell_box(width=16, height=12)
    cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
    spp = 3
    _, w, _ = oracle.render_path(sc, spp, cam)
    # interior pixels: samples of pixels (x-1..x, y-1..y) always cover (x, y)
    assert (w[1:, 1:] >= 4 * spp).all() and (w[1:, 1:] <= 9 * spp).all()
    assert w[0, 0] >= spp


def test_directlighting_furnace_is_le_times_one_plus_rho():
    """directlighting = emission + one unweighted light sample: inside the
    furnace every surface sees radiance Le from all directions, so the
    estimate converges to Le (1 + rho)."""
    for rho in (0.3, 0.7):
        sc = furnace_scene(rho=rho)
        cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
        rgb, w, _ = oracle.render_direct(sc, 32, cam)
        mean = float((rgb / w[..., None]).mean())
        assert abs(mean - (1.0 + rho)) < 0.02 * (1.0 + rho), (rho, mean)


def test_interior_generator():
    """The wavefront-at-scale scene family: triangle budget, 64 equal-area
    downward ceiling lights, inward room faces, the material mix present."""
    sc = scenes.interior(120_000, seed=3, width=32, height=18)
    assert abs(len(sc.tris) - 120_000) < 200
    lights = sc.meshes[sc.meshes["is_light"] == 1]
    assert len(lights) == 64 and (lights["num_tris"] == 2).all()
    for m in lights[:4]:
        a, b = sc.tris[m["first_tri"]].reshape(3, 3), sc.tris[m["first_tri"] + 1].reshape(3, 3)
        na, nb = np.cross(a[1] - a[0], a[2] - a[0]), np.cross(b[1] - b[0], b[2] - b[0])
        assert na[1] < 0 and nb[1] < 0 and abs(np.linalg.norm(na) - np.linalg.norm(nb)) < 1e-6
    centre = np.array([5.0, 2.0, 5.0])
    for m in sc.meshes[:6]:
        t = sc.tris[m["first_tri"]].reshape(3, 3).astype(np.float64)
        assert np.dot(np.cross(t[1] - t[0], t[2] - t[0]), centre - t.mean(0)) > 0
    assert set(np.unique(sc.meshes["material"])) == {scenes.MATTE, scenes.PLASTIC, scenes.MIRROR, scenes.GLASS}
    # textures on the floor, the far wall and some matte / plastic objects only
    tex = sc.meshes["texture"]
    assert len(sc.textures) == 4 and all(t.shape == (1024, 1024, 4) for t in sc.textures)
    assert tex[0] == 0 and tex[3] == 1 and tex.max() < 4
    assert np.isin(sc.meshes["material"][tex >= 0], [scenes.MATTE, scenes.PLASTIC]).all()
    assert 0.2 < (tex[70:] >= 0).mean() < 0.5
    assert (sc.meshes["kd"][(tex >= 0) & (sc.meshes["material"] == scenes.MATTE)] == 1).all()
    assert sc.uv.shape == (len(sc.tris), 6) and np.isfinite(sc.uv).all()
    # the texture draws leave the geometry and materials alone
    plain = scenes.interior(120_000, seed=3, width=32, height=18, textures=False)
    np.testing.assert_array_equal(plain.tris, sc.tris)
    np.testing.assert_array_equal(plain.meshes["material"], sc.meshes["material"])
    assert plain.uv is None and plain.textures == []


def test_vertex_normals_equal_to_geometric_change_nothing():
    """Shading with per-vertex normals that all equal the face normal gives
    the film of flat shading, for every integrator."""
    sc = scenes.cornell_box(seed=2, width=24, height=20, materials="mixed")
    cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
    t = sc.tris.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    flat = scenes.SceneArrays(sc.tris, sc.meshes, sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height,
                              sc.max_depth, normals=np.repeat(n, 3, axis=0).reshape(-1, 9),
                              mesh_normals=np.ones(len(sc.meshes), np.uint8))
    for fn in (oracle.render_path, oracle.render_direct):
        a, wa, _ = fn(sc, 2, cam)
        b, wb, _ = fn(flat, 2, cam)
        np.testing.assert_array_equal(wa, wb)
        np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)


def test_smooth_sphere_changes_its_shading_only():
    sc = scenes.cornell_box(seed=2, width=32, height=32)
    sm = scenes.cornell_box(seed=2, width=32, height=32, smooth=True)
    cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
    a, _, _ = oracle.render_direct(sc, 2, cam)
    b, _, _ = oracle.render_direct(sm, 2, cam)
    changed = np.any(a != b, axis=-1)
    assert 0 < changed.mean() < 0.2  # the sphere covers a small part of the image


def _bilinear64(t, u, v):
    """Independent float64 statement of a wrap-addressed bilinear read."""
    h, w = t.shape[:2]
    x, y = (u % 1.0) * w - 0.5, (v % 1.0) * h - 0.5
    x0, y0 = int(np.floor(x)), int(np.floor(y))
    a, b = x - x0, y - y0
    c = t[..., :3].astype(np.float64) / 255
    at = lambda i, j: c[j % h, i % w]
    return (1 - b) * ((1 - a) * at(x0, y0) + a * at(x0 + 1, y0)) + b * ((1 - a) * at(x0, y0 + 1) + a * at(x0 + 1, y0 + 1))


def test_texture_fetch_known_answers():
    """Texel centres return the texel exactly; half-way points average two
    texels exactly; the image wraps around on both axes (u = 0 averages the
    first and last columns) and in whole periods; random points agree with a
    float64 statement of the filter."""
    rng = np.random.default_rng(4)
    t = rng.integers(0, 256, (4, 8, 4)).astype(np.uint8)
    c = t[..., :3].astype(np.float32) / np.float32(255)
    for j in range(4):
        for i in range(8):
            u, v = (i + 0.5) / 8, (j + 0.5) / 4
            np.testing.assert_array_equal(oracle.tex_fetch(t, u, v), c[j, i])
            np.testing.assert_array_equal(oracle.tex_fetch(t, u + 3, v - 2), c[j, i])
    half = np.float32(0.5)
    np.testing.assert_array_equal(oracle.tex_fetch(t, 3 / 8, 0.125), half * c[0, 2] + half * c[0, 3])
    np.testing.assert_array_equal(oracle.tex_fetch(t, 0.0, 0.125), half * c[0, 7] + half * c[0, 0])
    np.testing.assert_array_equal(oracle.tex_fetch(t, 1 / 16, 0.0), half * c[3, 0] + half * c[0, 0])
    for _ in range(200):
        u, v = rng.uniform(-3, 3, 2).astype(np.float32)
        np.testing.assert_allclose(oracle.tex_fetch(t, u, v), _bilinear64(t, float(u), float(v)), atol=2e-6)
    # non-finite coordinates read at 0
    np.testing.assert_array_equal(oracle.tex_fetch(t, np.nan, np.inf), oracle.tex_fetch(t, 0.0, 0.0))


def _with_texture(sc, mesh, tex):
    meshes = sc.meshes.copy()
    meshes["texture"][mesh] = 0
    return scenes.SceneArrays(sc.tris, meshes, sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height,
                              sc.max_depth, uv=scenes.planar_uv(sc.tris), textures=[tex])


def test_constant_texture_equals_constant_colour():
    """A constant Kd texture T multiplies a matte colour (Kd T / 255) and
    replaces a plastic's diffuse colour (T / 255): the films equal those of
    the untextured scenes with those colours; mirror and glass ignore it."""
    sc = scenes.cornell_box(seed=2, width=24, height=20, materials="mixed")
    cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
    T = np.array([200, 60, 120, 255], np.uint8)
    tex = np.tile(T, (4, 4, 1))
    col = T[:3].astype(np.float32) / np.float32(255)
    for mesh in (0, 5, 6):   # matte floor, plastic tall box, mirror short box
        textured = _with_texture(sc, mesh, tex)
        meshes = sc.meshes.copy()
        if meshes["material"][mesh] == scenes.MATTE:
            meshes["kd"][mesh] = meshes["kd"][mesh] * col
        elif meshes["material"][mesh] == scenes.PLASTIC:
            meshes["kd"][mesh] = col
        plain = scenes.SceneArrays(sc.tris, meshes, sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height,
                                   sc.max_depth)
        for fn in (oracle.render_path, oracle.render_direct):
            a, wa, _ = fn(textured, 2, cam)
            b, wb, _ = fn(plain, 2, cam)
            np.testing.assert_array_equal(wa, wb)
            np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)
        a, _, _ = oracle.render_rlpath(textured, 2, cam, max_depth=2)
        b, _, _ = oracle.render_rlpath(plain, 2, cam, max_depth=2)
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-5)


def test_textures_change_only_textured_surfaces():
    sc = scenes.cornell_box(seed=2, width=32, height=32)
    tx = scenes.cornell_box(seed=2, width=32, height=32, textured=True)
    cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
    a, _, _ = oracle.render_direct(sc, 1, cam)
    b, _, _ = oracle.render_direct(tx, 1, cam)
    changed = np.any(a != b, axis=-1).mean()
    assert 0.2 < changed < 0.9


def test_crop_equals_the_full_film():
    """A crop traces only the samples that reach its pixels, with the full
    image's sample indices: its films equal that part of the full films,
    including the border, where pixels outside the image get no samples."""
    sc = scenes.cornell_box(seed=3, width=40, height=30, materials="mixed", textured=True)
    cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
    rgb, w, _ = oracle.render_path(sc, 4, cam)
    for x0, y0, x1, y1 in ((10, 7, 17, 12), (0, 0, 3, 2), (35, 26, 40, 30)):
        crgb, cw, _, smp = oracle.render_path(sc, 4, cam, crop=(x0, y0, x1, y1), samples=True)
        np.testing.assert_array_equal(crgb, rgb[y0:y1, x0:x1])
        np.testing.assert_array_equal(cw, w[y0:y1, x0:x1])
        # film positions lie within one pixel of their pixel's corner
        xs = np.arange(x0, x1)[None, :, None]
        assert ((smp[..., 3] >= xs) & (smp[..., 3] < xs + 1)).all()


def test_rlpath_furnace_is_le_over_one_minus_rho():
    """rlpath learns its light choice and its 128-bin direction distribution
    as it goes (Integrators/ReinforcementLearningPathTracing.cu); whatever it
    has learnt, the MIS weights of the light sample and of the Q-sampled hit
    sum to 1, so the furnace's Le / (1 - rho) still holds over many bounces."""
    for rho in (0.3, 0.6):
        sc = furnace_scene(rho=rho)
        cam = look_at(sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height)
        rgb, w, _ = oracle.render_rlpath(sc, 32, cam)
        expect = 1.0 / (1.0 - rho)
        mean = float((rgb / w[..., None]).mean())
        assert abs(mean - expect) < 0.01 * expect, (rho, mean, expect)
