"""GPU vs oracle at the sizes of the benchmark workloads.

* LBVH + traversal: 2^20 Gaussian-mixture triangles, 3 treelet rounds,
  2^18 closest-hit and 2^18 shadow rays; nodes bit for bit, hit ids and
  t/u/v bit for bit, visibility exact, no stack overflow.
* Wavefront path tracing at scale: the 2M-triangle procedural interior
  (displaced icospheres and boxes, 64 quad lights, matte / plastic / mirror /
  glass, 1024^2 textures) at 1920x1080, 2 spp, max depth 16, 3 treelet
  rounds; 64x64 crops centred on a mirror, a glass, a plastic and a textured
  object, on the textured floor and on a light are compared with the
  oracle's crop renderer bit for bit (weights and sums).
* rlpath on the same scene: deterministic mode at q_grid 64 and 32 (the
  reference's 32^3), 96x54 pixels, 2 spp, depth 16, bit for bit.
The interior's BVH is built once by the oracle (3 treelet rounds) and
compared with the GPU's before any render uses it."""
import numpy as np
import pytest

import oracle
from cavalry_amd import BVH, render as R, scenes
from cavalry_amd.camera import look_at
from helpers import assert_nodes_equal, bits

pytestmark = pytest.mark.gpu

MATTE, MIRROR, GLASS, PLASTIC = 0, 1, 4, 5


def test_lbvh_and_trace_2p20(ctx):
    n, nr = 1 << 20, 1 << 18
    print(f"[config1] n={n} rounds=3 rays={nr}+{nr}", flush=True)
    tris = scenes.gmm_triangles(n, seed=20)
    codes, order, nd = oracle.lbvh_build(tris, 30)
    oracle.treelet_optimize(nd, 3)
    bvh = BVH.build(ctx, tris, optimization_rounds=3, morton_bits=30)
    ex = bvh.export()
    np.testing.assert_array_equal(ex["codes"], codes)
    np.testing.assert_array_equal(ex["order"], order)
    assert_nodes_equal(ex, nd, "n=2^20 rounds=3: ")
    depth = bvh.max_depth()
    print(f"[config1] nodes bit-exact; deepest leaf {depth}", flush=True)
    assert depth <= 64  # no traversal can overflow the reference's 64-entry stack

    rays = scenes.random_rays(nr, seed=21)
    ref = oracle.trace_closest(nd, tris, rays)
    got = bvh.intersect(rays)
    np.testing.assert_array_equal(got["prim"], ref["prim"])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(bits(got[k]), bits(ref[k]), err_msg=k)
    srays = scenes.random_rays(nr, seed=22, stream=2)
    lim = scenes.shadow_limits(nr, seed=23)
    np.testing.assert_array_equal(bvh.test_visibility(srays, lim), oracle.trace_occluded(nd, tris, srays, lim))
    c = bvh.trace_counters(reset=True)
    print(f"[config1] hits={int((ref['prim'] >= 0).sum())} counters={c}", flush=True)
    assert c["overflow_closest"] == 0 and c["overflow_occluded"] == 0
    assert 0 < c["max_stack_closest"] <= depth and c["max_stack_occluded"] <= depth
    bvh.close()


# ---------------------------------------------------------------- interior
def _project(cam, pts):
    """Pixel coordinates of world points (the inverse of the camera's ray
    generation: film point (x - W/2, H/2 - y, depth) in camera space)."""
    m = np.array(cam.cam_to_world, np.float64).reshape(4, 4).T  # column-major
    inv = np.linalg.inv(m)
    c = (inv[:3, :3] @ np.asarray(pts, np.float64).T + inv[:3, 3:4]).T
    x = cam.width / 2 + c[:, 0] * cam.depth / c[:, 2]
    y = cam.height / 2 - c[:, 1] * cam.depth / c[:, 2]
    return x, y, c[:, 2]


def _pixel_rays(cam, x0, y0, x1, y1):
    """Camera rays through the pixel centres of a crop."""
    m = np.array(cam.cam_to_world, np.float64).reshape(4, 4).T
    ys, xs = np.mgrid[y0:y1, x0:x1]
    p = np.stack([xs + 0.5 - cam.width / 2, cam.height / 2 - (ys + 0.5), np.full(xs.shape, cam.depth)], -1)
    d = p.reshape(-1, 3) @ m[:3, :3].T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(m[:3, 3], d.shape)
    return np.ascontiguousarray(np.concatenate([o, d], 1), np.float32)


@pytest.fixture(scope="module")
def interior2m(ctx):
    print("[interior] generating 2M triangles, oracle BVH with 3 treelet rounds", flush=True)
    arr = scenes.interior(2_000_000, seed=1, width=1920, height=1080, max_depth=16)
    tris = np.ascontiguousarray(arr.tris, np.float32)
    _, _, nd = oracle.lbvh_build(tris, 30)
    oracle.treelet_optimize(nd, 3)
    sc = R.Scene(ctx, arr, bvh_rounds=3)
    assert_nodes_equal(sc.bvh().export(), nd, "interior 2M rounds=3: ")
    print(f"[interior] {len(tris)} triangles, {len(arr.meshes)} meshes, BVH bit-exact, "
          f"deepest leaf {sc.bvh().max_depth()}", flush=True)
    yield arr, sc, nd
    sc.close()


def _crops(arr, cam, nd, size=64, stride=4):
    """For each kind of first hit (mirror, glass, untextured plastic,
    textured surface, light), the size x size crop that shows the most of it,
    found by tracing pixel-centre rays on a grid of `stride` pixels; plus the
    middle of the floor. Returns [(name, x0, y0, x1, y1)]."""
    m = arr.meshes
    tris = np.ascontiguousarray(arr.tris, np.float32)
    W, H = cam.width, cam.height
    gh, gw = H // stride, W // stride
    rays = _pixel_rays(cam, 0, 0, W, H).reshape(H, W, 6)[::stride, ::stride][:gh, :gw].reshape(-1, 6)
    prim = oracle.trace_closest(nd, tris, np.ascontiguousarray(rays))["prim"].reshape(gh, gw)
    mesh = np.where(prim >= 0, arr.tri_mesh[np.maximum(prim, 0)], -1)
    mm = m[np.maximum(mesh, 0)]
    hit = mesh >= 0
    kinds = {"mirror": hit & (mm["material"] == MIRROR), "glass": hit & (mm["material"] == GLASS),
             "plastic": hit & (mm["material"] == PLASTIC) & (mm["texture"] < 0),
             "textured": hit & (mm["texture"] >= 0), "light": hit & (mm["is_light"] != 0)}
    k = size // stride
    out = []
    for name, mask in kinds.items():
        c = np.zeros((gh + 1, gw + 1), np.int64)
        c[1:, 1:] = mask.cumsum(0).cumsum(1)
        win = c[k:, k:] - c[:-k, k:] - c[k:, :-k] + c[:-k, :-k]  # count in each k x k window
        iy, ix = np.unravel_index(int(np.argmax(win)), win.shape)
        assert win[iy, ix] >= k * k // 10, f"no crop shows much {name} ({win[iy, ix]} of {k * k})"
        out.append((name, ix * stride, iy * stride, ix * stride + size, iy * stride + size))
    out.append(("floor", W // 2 - size // 2, H - size - 8, W // 2 + size // 2, H - 8))
    return out


def test_interior_path_crops_match_oracle(ctx, interior2m):
    arr, sc, nd = interior2m
    cam = look_at(arr.eye, arr.center, arr.up, arr.fov, 1920, 1080)
    print("[config2] GPU path 1920x1080 2 spp depth 16", flush=True)
    film, st = R.render(ctx, sc, 2, camera=cam, max_depth=16)
    assert st["closest_rays"] > 1920 * 1080 * 2
    for name, x0, y0, x1, y1 in _crops(arr, cam, nd):
        rgb, w, _ = oracle.render_path(arr, 2, cam, crop=(x0, y0, x1, y1), max_depth=16, bvh=nd)
        g_rgb, g_w = film.rgb[y0:y1, x0:x1], film.weight[y0:y1, x0:x1]
        diff = int(np.count_nonzero((bits(g_rgb) != bits(rgb)).any(-1)))
        print(f"[config2] crop {name} ({x0},{y0})-({x1},{y1}): mean={float(rgb.sum() / w.sum() / 3):.4f} "
              f"pixels differing={diff}", flush=True)
        np.testing.assert_array_equal(g_w, w, err_msg=name)
        np.testing.assert_array_equal(bits(g_rgb), bits(rgb), err_msg=name)


@pytest.mark.parametrize("q_grid", [64, 32])
def test_interior_rlpath_deterministic_matches_oracle(ctx, interior2m, q_grid):
    arr, sc, nd = interior2m
    cam = look_at(arr.eye, arr.center, arr.up, arr.fov, 96, 54)
    print(f"[config3] rlpath q_grid={q_grid} 96x54 2 spp depth 16 deterministic", flush=True)
    film, st = R.render(ctx, sc, 2, integrator="rlpath", camera=cam, max_depth=16, q_grid=q_grid,
                        deterministic=True)
    rgb, w, rays = oracle.render_rlpath(arr, 2, cam, q_grid=q_grid, max_depth=16, bvh=nd)
    diff = int(np.count_nonzero((bits(film.rgb) != bits(rgb)).any(-1)))
    print(f"[config3] rays gpu={st['closest_rays'] + st['shadow_rays']} oracle={rays} pixels differing={diff}",
          flush=True)
    assert st["closest_rays"] + st["shadow_rays"] == rays
    np.testing.assert_array_equal(film.weight, w)
    np.testing.assert_array_equal(bits(film.rgb), bits(rgb))


def test_interior_limit_cull_keeps_the_film(ctx, interior2m):
    """The whole 1920x1080 config-2 film (2 spp, depth 16): any-hit with the
    limit cull (the default) and along the whole line, the reference's
    visiting order, give the same film bit for bit (every shadow ray's
    visibility is the same)."""
    from cavalry_amd import tune
    arr, sc, _ = interior2m
    cam = look_at(arr.eye, arr.center, arr.up, arr.fov, 1920, 1080)
    print("[config2] path 1920x1080 2 spp depth 16, limit cull vs whole line", flush=True)
    saved = tune.get_launch(ctx, tune.WAVEFRONT)
    try:
        a, _ = R.render(ctx, sc, 2, camera=cam, max_depth=16)
        tune.set_launch(ctx, tune.WAVEFRONT, dict(saved, any_hit_whole_line=1))
        b, _ = R.render(ctx, sc, 2, camera=cam, max_depth=16)
    finally:
        tune.set_launch(ctx, tune.WAVEFRONT, saved)
    np.testing.assert_array_equal(a.weight, b.weight)
    np.testing.assert_array_equal(bits(a.rgb), bits(b.rgb))
