"""Scenes loaded from .pbrt files render exactly like the same scenes built
from arrays, through the Python API and through the command-line renderer."""
import os
import re
import subprocess
import zlib
import struct

import numpy as np
import pytest

from cavalry_amd import pbrt, render as R, scenes
from cavalry_amd.camera import for_scene

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cavalry_amd", "_lib", "cavalry")


def read_png_rgb(path):
    data = open(path, "rb").read()
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            w, h = struct.unpack(">II", body[:8])
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    rows = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    return rows[:, 1:].reshape(h, w, 3)


@pytest.mark.parametrize("integrator", ["path", "rlpath"])
def test_pbrt_scene_renders_like_arrays(ctx, tmp_path, integrator):
    print(f"[pbrt] {integrator}", flush=True)
    sc = scenes.cornell_box(seed=3, width=48, height=40, materials="mixed")
    p = str(tmp_path / "c.pbrt")
    pbrt.write(sc, p, spp=4, integrator=integrator)
    a = R.Scene(ctx, sc, bvh_rounds=3, morton_bits=63)
    film_a, _ = R.render(ctx, a, 4, integrator=integrator, camera=for_scene(sc), max_depth=1)
    b, st = R.Scene.from_pbrt(ctx, p)
    assert st.spp == 4 and st.bvh_opt_rounds == 3 and st.morton_bits == 63
    film_b, _ = R.render(ctx, b, st.spp, integrator=integrator, max_depth=1)
    np.testing.assert_array_equal(film_b.weight, film_a.weight)
    np.testing.assert_array_equal(film_b.rgb, film_a.rgb)
    a.close()
    b.close()


@pytest.mark.parametrize("integrator", ["path", "directlighting"])
def test_pbrt_vertex_normals_render_like_arrays(ctx, tmp_path, integrator):
    """A smooth-shaded scene from a .pbrt file carries its vertex normals to
    the device: it renders exactly like the array scene built with the
    parsed normals, and differently from the same scene without them."""
    print(f"[pbrt] normals {integrator}", flush=True)
    # matte: with the mixed materials the smooth sphere is glass, whose light
    # samples contribute nothing to directlighting, normals or not
    sc = scenes.cornell_box(seed=6, width=48, height=40, materials="matte", smooth=True)
    p = str(tmp_path / "s.pbrt")
    pbrt.write(sc, p, spp=4, integrator=integrator)
    normals, flags = pbrt.parse_normals(p)
    arr = scenes.SceneArrays(sc.tris, sc.meshes, sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height,
                             sc.max_depth, normals=normals, mesh_normals=flags)
    flat = scenes.SceneArrays(sc.tris, sc.meshes, sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height,
                              sc.max_depth)
    b, st = R.Scene.from_pbrt(ctx, p)
    film_b, _ = R.render(ctx, b, st.spp, integrator=integrator, max_depth=2)
    for arrays, same in ((arr, True), (flat, False)):
        a = R.Scene(ctx, arrays, bvh_rounds=st.bvh_opt_rounds, morton_bits=st.morton_bits)
        film_a, _ = R.render(ctx, a, st.spp, integrator=integrator, camera=for_scene(sc), max_depth=2)
        np.testing.assert_array_equal(film_b.weight, film_a.weight)
        assert np.array_equal(film_b.rgb, film_a.rgb) == same
        a.close()
    b.close()


@pytest.mark.parametrize("integrator", ["path", "rlpath"])
def test_pbrt_textures_render_like_arrays(ctx, tmp_path, integrator):
    """A textured scene from a .pbrt file (PNG textures, "float uv", Kd
    textures on matte and plastic) renders exactly like the array scene built
    from the parsed meshes, texture coordinates and textures."""
    print(f"[pbrt] textures {integrator}", flush=True)
    sc = scenes.cornell_box(seed=8, width=48, height=40, materials="mixed", textured=True)
    p = str(tmp_path / "t.pbrt")
    pbrt.write(sc, p, spp=4, integrator=integrator)
    tris, meshes, st, _ = pbrt.parse(p)
    uv, textures = pbrt.parse_textures(p)
    arr = scenes.SceneArrays(tris, meshes, sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height, sc.max_depth,
                             uv=uv, textures=textures)
    b, _ = R.Scene.from_pbrt(ctx, p)
    film_b, _ = R.render(ctx, b, st.spp, integrator=integrator, max_depth=1)
    a = R.Scene(ctx, arr, bvh_rounds=st.bvh_opt_rounds, morton_bits=st.morton_bits)
    film_a, _ = R.render(ctx, a, st.spp, integrator=integrator, camera=for_scene(sc), max_depth=1)
    np.testing.assert_array_equal(film_b.weight, film_a.weight)
    np.testing.assert_array_equal(film_b.rgb, film_a.rgb)
    a.close()
    b.close()


@pytest.mark.parametrize("integrator", ["path", "directlighting"])
def test_pbrt_environment_renders_like_arrays(ctx, tmp_path, integrator):
    """A scene with a rotated, mapped "infinite" light and a black light
    read from a .pbrt file renders exactly like the array scene built from
    the parsed light list."""
    print(f"[pbrt] environment {integrator}", flush=True)
    sc = scenes.cornell_box(seed=9, width=48, height=40, materials="mixed", env=True)
    p = str(tmp_path / "e.pbrt")
    pbrt.write(sc, p, spp=4, integrator=integrator)
    tris, meshes, st, _ = pbrt.parse(p)
    uv, textures = pbrt.parse_textures(p)
    envs, order, background = pbrt.parse_lights(p)
    arr = scenes.SceneArrays(tris, meshes, sc.eye, sc.center, sc.up, sc.fov, sc.width, sc.height, sc.max_depth,
                             uv=uv, textures=textures, env_lights=envs, light_order=order,
                             env_background=background)
    b, _ = R.Scene.from_pbrt(ctx, p)
    film_b, _ = R.render(ctx, b, st.spp, integrator=integrator, max_depth=3)
    a = R.Scene(ctx, arr, bvh_rounds=st.bvh_opt_rounds, morton_bits=st.morton_bits)
    film_a, _ = R.render(ctx, a, st.spp, integrator=integrator, camera=for_scene(sc), max_depth=3)
    np.testing.assert_array_equal(film_b.weight, film_a.weight)
    np.testing.assert_array_equal(film_b.rgb, film_a.rgb)
    a.close()
    b.close()


def test_cli_renders_png(ctx, tmp_path):
    print("[pbrt] CLI", flush=True)
    sc = scenes.cornell_box(seed=4, width=40, height=32)
    p = str(tmp_path / "c.pbrt")
    pbrt.write(sc, p, spp=2)
    out = str(tmp_path / "out.png")
    pfm = str(tmp_path / "out.pfm")
    r = subprocess.run([CLI, "--input", p, "--output", out, "--spp", "3", "--bvh-opt-rounds", "1", "--pfm", pfm,
                        "--unknown-flag"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert re.search(r"^Rendering took \d+ms$", r.stdout, re.M)
    s, st = R.Scene.from_pbrt(ctx, p, spp=3, bvh_opt_rounds=1)
    film, _ = R.render(ctx, s, 3)
    np.testing.assert_array_equal(read_png_rgb(out), film.to_srgb8())
    np.testing.assert_array_equal(R.read_pfm(pfm), film.image)
    s.close()
