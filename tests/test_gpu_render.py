"""Wavefront `path` renderer on the GPU vs the CPU oracle on the same scene,
camera and Halton sequence. The film must match to per-pixel relative MSE
<= 1e-4; in practice it is bit-identical apart from the rare samples whose
float cos/sin (GGX near-normal case) differs between libm and the device."""
import numpy as np
import pytest

import oracle
from cavalry_amd import Context, render as R, scenes
from cavalry_amd.camera import for_scene

pytestmark = pytest.mark.gpu

REL_MSE_TOL = 1e-4


def rel_mse(a, b):
    """mean over pixels and channels of (a - b)^2 / (b^2 + 1e-2)."""
    return float(np.mean((a - b) ** 2 / (b * b + 1e-2)))


@pytest.mark.parametrize("materials,rounds", [("matte", 0), ("matte", 3), ("mixed", 1)])
def test_path_matches_oracle(ctx, materials, rounds):
    print(f"[render] cornell {materials} rounds={rounds}", flush=True)
    sc = scenes.cornell_box(seed=3, width=64, height=48, materials=materials)
    cam = for_scene(sc)
    spp = 4
    rgb, w, rays = oracle.render_path(sc, spp, cam, bvh_rounds=rounds)
    scene = R.Scene(ctx, sc, bvh_rounds=rounds)
    film, st = R.render(ctx, scene, spp, camera=cam)
    np.testing.assert_array_equal(film.weight, w)
    a, b = film.image, rgb / np.maximum(w, 1e-30)[..., None]
    err = rel_mse(a, b)
    same = float(np.mean(np.all(film.rgb.view(np.uint32) == rgb.view(np.uint32), axis=-1)))
    print(f"[render] rel_mse={err:.3e} identical pixels={same:.4f} rays={st['closest_rays'] + st['shadow_rays']} oracle rays={rays}", flush=True)
    assert err <= REL_MSE_TOL
    assert same > 0.99
    assert st["closest_rays"] > 0 and st["shadow_rays"] > 0
    scene.close()


def test_render_errors(ctx):
    from cavalry_amd._lib import CavError
    sc = scenes.cornell_box(width=8, height=8)
    m = sc.meshes.copy()
    m["is_light"] = 0
    bad = scenes.SceneArrays(sc.tris, m, sc.eye, sc.center, sc.up, sc.fov, 8, 8)
    with pytest.raises(CavError):
        R.Scene(ctx, bad)
    scene = R.Scene(ctx, sc)
    with pytest.raises(CavError):
        R.render(ctx, scene, 0)
    with pytest.raises(CavError):
        R.render(ctx, scene, 1, max_depth=200)


def test_path_rounds_equal_single_round_oracle(ctx):
    """spp rounds (sppBySpp order) keep each sample's Halton index, so the
    film does not depend on the split."""
    print("[render] cornell matte spp=5 in rounds of 2", flush=True)
    sc = scenes.cornell_box(seed=5, width=40, height=30)
    cam = for_scene(sc)
    rgb, w, rays = oracle.render_path(sc, 5, cam)
    scene = R.Scene(ctx, sc)
    film, st = R.render(ctx, scene, 5, camera=cam, spp_per_round=2)
    assert st["rounds"] == 3
    np.testing.assert_array_equal(film.weight, w)
    np.testing.assert_array_equal(film.rgb, rgb)
    assert st["closest_rays"] + st["shadow_rays"] == rays
    film1, st1 = R.render(ctx, scene, 5, camera=cam)
    assert st1["rounds"] == 1
    np.testing.assert_array_equal(film1.rgb, rgb)
    scene.close()


def test_film_f64_sums(ctx):
    """FILM_F64 adds the same samples in the same order in binary64: the
    weights (small integers) are identical, the sums agree with the binary32
    film to the rounding of at most 9 x spp binary32 additions, and the float
    film after a later plain render is the oracle's again (the flag does not
    stick to the context's buffers)."""
    print("[render] cornell matte spp=64 film_f64", flush=True)
    sc = scenes.cornell_box(seed=6, width=40, height=30)
    cam = for_scene(sc)
    scene = R.Scene(ctx, sc)
    f32, _ = R.render(ctx, scene, 64, camera=cam)
    f64, _ = R.render(ctx, scene, 64, camera=cam, film_f64=True)
    np.testing.assert_array_equal(f64.weight, f32.weight)
    np.testing.assert_allclose(f64.rgb, f32.rgb, rtol=9 * 64 * 2 ** -24, atol=1e-6)
    rgb, w, _ = oracle.render_path(sc, 4, cam)
    again, _ = R.render(ctx, scene, 4, camera=cam)
    np.testing.assert_array_equal(again.weight, w)
    assert rel_mse(again.image, rgb / np.maximum(w, 1e-30)[..., None]) <= REL_MSE_TOL
    scene.close()


@pytest.mark.parametrize("materials", ["matte", "mixed"])
def test_directlighting_matches_oracle(ctx, materials):
    print(f"[render] directlighting {materials}", flush=True)
    sc = scenes.cornell_box(seed=7, width=48, height=40, materials=materials)
    cam = for_scene(sc)
    rgb, w, rays = oracle.render_direct(sc, 4, cam)
    scene = R.Scene(ctx, sc)
    film, st = R.render(ctx, scene, 4, integrator="directlighting", camera=cam)
    np.testing.assert_array_equal(film.weight, w)
    np.testing.assert_array_equal(film.rgb, rgb)
    assert st["closest_rays"] + st["shadow_rays"] == rays
    scene.close()


@pytest.mark.parametrize("integrator,depth", [("path", 8), ("directlighting", 1), ("rlpath", 1)])
def test_material_sort_keeps_the_film(ctx, integrator, depth):
    """Shading hits in sortLightingQueue's order changes only the speed."""
    print(f"[render] material sort {integrator}", flush=True)
    sc = scenes.cornell_box(seed=8, width=48, height=40, materials="mixed")
    cam = for_scene(sc)
    scene = R.Scene(ctx, sc)
    a, sa = R.render(ctx, scene, 4, integrator=integrator, camera=cam, max_depth=depth)
    b, sb = R.render(ctx, scene, 4, integrator=integrator, camera=cam, max_depth=depth, sort_by_material=True)
    np.testing.assert_array_equal(a.rgb, b.rgb)
    np.testing.assert_array_equal(a.weight, b.weight)
    assert sa["closest_rays"] == sb["closest_rays"] and sa["shadow_rays"] == sb["shadow_rays"]
    scene.close()


# (width, height, spp): one histogram tile of the material sort (at most
# 131072 hits), two tiles, and four (the carry across tiles of its scan)
SORT_SIZES = [(48, 40, 4), (256, 256, 4), (384, 384, 3)]


@pytest.mark.parametrize("w,h,spp", SORT_SIZES)
def test_material_sorted_path_matches_oracle(ctx, w, h, spp):
    """path with the material-sorted shading order (sortLightingQueue,
    PathTracingIntegrator.cu:50-92) against the oracle, bit for bit."""
    print(f"[render] material-sorted path {w}x{h}x{spp}", flush=True)
    sc = scenes.cornell_box(seed=9, width=w, height=h, materials="mixed")
    cam = for_scene(sc)
    rgb, wt, rays = oracle.render_path(sc, spp, cam)
    scene = R.Scene(ctx, sc)
    film, st = R.render(ctx, scene, spp, camera=cam, sort_by_material=True)
    np.testing.assert_array_equal(film.weight, wt)
    np.testing.assert_array_equal(film.rgb.view(np.uint32), rgb.view(np.uint32))
    assert st["closest_rays"] + st["shadow_rays"] == rays
    scene.close()


@pytest.mark.parametrize("w,h,spp", SORT_SIZES)
def test_entry_sorted_rlpath_deterministic(ctx, w, h, spp):
    """Deterministic rlpath with the (material, Q entry) shading order of
    the reference's rlpath sortLightingQueue (RL :226-272) equals the
    unsorted render bit for bit, and the oracle's."""
    print(f"[rlpath] entry-sorted deterministic {w}x{h}x{spp}", flush=True)
    sc = scenes.cornell_box(seed=10, width=w, height=h, materials="mixed")
    cam = for_scene(sc)
    scene = R.Scene(ctx, sc)
    a, sa = R.render(ctx, scene, spp, integrator="rlpath", camera=cam, deterministic=True)
    b, sb = R.render(ctx, scene, spp, integrator="rlpath", camera=cam, deterministic=True, sort_by_material=True)
    np.testing.assert_array_equal(b.weight, a.weight)
    np.testing.assert_array_equal(b.rgb.view(np.uint32), a.rgb.view(np.uint32))
    assert sa["closest_rays"] == sb["closest_rays"] and sa["shadow_rays"] == sb["shadow_rays"]
    rgb, wt, rays = oracle.render_rlpath(sc, spp, cam)
    np.testing.assert_array_equal(b.weight, wt)
    np.testing.assert_array_equal(b.rgb.view(np.uint32), rgb.view(np.uint32))
    assert sb["closest_rays"] + sb["shadow_rays"] == rays
    scene.close()


@pytest.mark.parametrize("integrator", ["path", "directlighting", "rlpath"])
def test_vertex_normals_match_oracle(ctx, integrator):
    """Smooth-shaded sphere (interpolated per-vertex normals at hits and at
    light samples): films equal the oracle's."""
    print(f"[render] vertex normals {integrator}", flush=True)
    sc = scenes.cornell_box(seed=9, width=48, height=40, materials="mixed", smooth=True)
    cam = for_scene(sc)
    depth = 1 if integrator == "rlpath" else sc.max_depth
    fn = {"path": oracle.render_path, "directlighting": oracle.render_direct}.get(integrator)
    if fn is None:
        rgb, w, rays = oracle.render_rlpath(sc, 4, cam, max_depth=1)
    else:
        rgb, w, rays = fn(sc, 4, cam)
    scene = R.Scene(ctx, sc)
    film, st = R.render(ctx, scene, 4, integrator=integrator, camera=cam, max_depth=depth)
    np.testing.assert_array_equal(film.weight, w)
    np.testing.assert_array_equal(film.rgb, rgb)
    assert st["closest_rays"] + st["shadow_rays"] == rays
    scene.close()


@pytest.mark.parametrize("integrator", ["path", "directlighting", "rlpath"])
def test_textures_match_oracle(ctx, integrator):
    """Kd textures (checker floor, striped wall, marble plastic box, planar
    and spherical UVs, bilinear wrap-addressed reads) together with vertex
    normals: films equal the oracle's."""
    print(f"[render] textures {integrator}", flush=True)
    sc = scenes.cornell_box(seed=10, width=48, height=40, materials="mixed", smooth=True, textured=True)
    cam = for_scene(sc)
    depth = 1 if integrator == "rlpath" else sc.max_depth
    if integrator == "rlpath":
        rgb, w, rays = oracle.render_rlpath(sc, 4, cam, max_depth=1)
    else:
        rgb, w, rays = {"path": oracle.render_path, "directlighting": oracle.render_direct}[integrator](sc, 4, cam)
    scene = R.Scene(ctx, sc)
    film, st = R.render(ctx, scene, 4, integrator=integrator, camera=cam, max_depth=depth)
    np.testing.assert_array_equal(film.weight, w)
    np.testing.assert_array_equal(film.rgb, rgb)
    assert st["closest_rays"] + st["shadow_rays"] == rays
    scene.close()


@pytest.mark.parametrize("integrator", ["path", "directlighting", "rlpath"])
def test_metal_substrate_match_oracle(ctx, integrator):
    """Metal (copper, and almost-delta alpha 0.04 that rlpath samples by the
    BSDF) and anisotropic substrate with a Kd texture: films equal the
    oracle's."""
    print(f"[render] metal/substrate {integrator}", flush=True)
    sc = scenes.cornell_box(seed=11, width=48, height=40, materials="metal", smooth=True, textured=True)
    cam = for_scene(sc)
    depth = 1 if integrator == "rlpath" else sc.max_depth
    if integrator == "rlpath":
        rgb, w, rays = oracle.render_rlpath(sc, 4, cam, max_depth=1)
    else:
        rgb, w, rays = {"path": oracle.render_path, "directlighting": oracle.render_direct}[integrator](sc, 4, cam)
    scene = R.Scene(ctx, sc)
    film, st = R.render(ctx, scene, 4, integrator=integrator, camera=cam, max_depth=depth)
    np.testing.assert_array_equal(film.weight, w)
    np.testing.assert_array_equal(film.rgb, rgb)
    assert st["closest_rays"] + st["shadow_rays"] == rays
    scene.close()


@pytest.mark.parametrize("integrator", ["path", "directlighting", "rlpath"])
@pytest.mark.parametrize("kind", ["cornell", "sky"])
def test_environment_lights_match_oracle(ctx, integrator, kind):
    """Environment lights: a rotated textured sky seen through the open front
    of the Cornell box plus a black light, in a custom light order; and a
    floor under a constant sky with no area light at all (the light sample
    then takes 2 Halton dimensions instead of 3). Films equal the oracle's."""
    print(f"[render] environment {kind} {integrator}", flush=True)
    if kind == "cornell":
        sc = scenes.cornell_box(seed=12, width=48, height=40, materials="mixed", smooth=True, textured=True,
                                env=True)
    else:
        sc = scenes.sky_plane(width=48, height=40, albedo=0.6, sky=1.3, max_depth=4, black_lights=1)
    cam = for_scene(sc)
    depth = 1 if integrator == "rlpath" else sc.max_depth
    if integrator == "rlpath":
        rgb, w, rays = oracle.render_rlpath(sc, 4, cam, max_depth=1)
    else:
        rgb, w, rays = {"path": oracle.render_path, "directlighting": oracle.render_direct}[integrator](sc, 4, cam)
    scene = R.Scene(ctx, sc)
    film, st = R.render(ctx, scene, 4, integrator=integrator, camera=cam, max_depth=depth)
    np.testing.assert_array_equal(film.weight, w)
    np.testing.assert_array_equal(film.rgb, rgb)
    assert st["closest_rays"] + st["shadow_rays"] == rays
    scene.close()


@pytest.mark.parametrize("integrator", ["path", "directlighting", "rlpath"])
@pytest.mark.parametrize("kind", ["cornell_sphere", "cornell_both", "plane"])
def test_sphere_lights_match_oracle(ctx, integrator, kind):
    """Sphere lights sampled within the cone they subtend, and the cone pdf
    in the MIS weight of BSDF rays that hit them: the Cornell box lit by a
    sphere (mixed materials, the sphere's own diffuse surface hit and shaded
    too), by the sphere and the ceiling quad, and a floor under a large
    sphere. Films equal the oracle's."""
    print(f"[render] sphere lights {kind} {integrator}", flush=True)
    if kind == "plane":
        sc = scenes.sphere_light_plane(width=48, height=40, radius=0.7, max_depth=4, light_kd=0.5)
    else:
        sc = scenes.cornell_box(seed=13, width=48, height=40, materials="mixed",
                                lights="sphere" if kind == "cornell_sphere" else "both")
    cam = for_scene(sc)
    depth = 1 if integrator == "rlpath" else sc.max_depth
    if integrator == "rlpath":
        rgb, w, rays = oracle.render_rlpath(sc, 4, cam, max_depth=1)
    else:
        rgb, w, rays = {"path": oracle.render_path, "directlighting": oracle.render_direct}[integrator](sc, 4, cam)
    scene = R.Scene(ctx, sc)
    film, st = R.render(ctx, scene, 4, integrator=integrator, camera=cam, max_depth=depth)
    np.testing.assert_array_equal(film.weight, w)
    np.testing.assert_array_equal(film.rgb, rgb)
    assert st["closest_rays"] + st["shadow_rays"] == rays
    scene.close()


def test_environment_errors(ctx):
    from cavalry_amd._lib import CavError
    print("[render] environment errors", flush=True)
    for order in ([-1, 8], [-1, 8, -2, -2], [-1, 3, -2], [-3, 8, -2]):
        sc = scenes.cornell_box(seed=1, width=16, height=16, env=True)
        sc.light_order = np.int32(order)
        with pytest.raises(CavError, match="light"):
            R.Scene(ctx, sc)
    sc = scenes.cornell_box(seed=1, width=16, height=16, env=True)
    sc.env_background = 2
    with pytest.raises(CavError, match="env_background"):
        R.Scene(ctx, sc)
    sc = scenes.cornell_box(seed=1, width=16, height=16, env=True)
    sc.env_lights["texture"][0] = 5
    with pytest.raises(CavError, match="texture"):
        R.Scene(ctx, sc)


def test_texture_errors(ctx):
    from cavalry_amd._lib import CavError
    print("[render] texture errors", flush=True)
    sc = scenes.cornell_box(seed=1, width=16, height=16, textured=True)
    sc.meshes["texture"][0] = 7  # only 3 textures
    with pytest.raises(CavError, match="texture"):
        R.Scene(ctx, sc)
    sc = scenes.cornell_box(seed=1, width=16, height=16, textured=True)
    sc.textures[1] = np.zeros((0, 4, 4), np.uint8)
    with pytest.raises(CavError, match="texture"):
        R.Scene(ctx, sc)
