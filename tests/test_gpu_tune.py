"""Traversal launch shapes on the GPU: every shape gives the same hits (the
node order of each ray is fixed), the C ABI checks its arguments, and the
autotuner picks a valid shape, caches it and reuses the cache."""
import numpy as np
import pytest
import torch

from cavalry_amd import BVH, render as R, scenes, tune
from cavalry_amd._lib import CavError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def traced(ctx):
    tris = scenes.gmm_triangles(1 << 16, seed=4)
    rays = scenes.random_rays(1 << 15, seed=5)
    srays = scenes.random_rays(1 << 15, seed=6, stream=2)
    lim = scenes.shadow_limits(1 << 15, seed=7)
    bvh = BVH.build(ctx, tris, optimization_rounds=2)
    yield bvh, rays, srays, lim
    bvh.close()


def test_every_shape_gives_the_same_hits(ctx, traced):
    bvh, rays, srays, lim = traced
    print("[tune] shapes", flush=True)
    saved = tune.get_launch(ctx, tune.STANDALONE)
    ref = None
    for size, per_cu in ((256, 1), (256, 3), (256, 0), (128, 5), (64, 12), (64, 0)):
        for chunk in (64, 1024):
            for lds in (32, 48):
                tune.set_launch(ctx, tune.STANDALONE, {"workgroups_per_cu": per_cu, "ray_chunk": chunk,
                                                       "lds_stack": lds, "workgroup_size": size})
                h = bvh.intersect(rays)
                v = bvh.test_visibility(srays, lim)
                got = (h["prim"].copy(), h["t"].copy(), v.copy())
                if ref is None:
                    ref = got
                for a, b in zip(got, ref):
                    np.testing.assert_array_equal(a, b)
    tune.set_launch(ctx, tune.STANDALONE, saved)


def test_launch_arguments_checked(ctx):
    with pytest.raises(CavError, match="lds_stack"):
        tune.set_launch(ctx, tune.STANDALONE, {"workgroups_per_cu": 2, "ray_chunk": 256, "lds_stack": 40})
    with pytest.raises(CavError, match="ray_chunk"):
        tune.set_launch(ctx, tune.WAVEFRONT, {"workgroups_per_cu": 2, "ray_chunk": 8, "lds_stack": 32})
    with pytest.raises(CavError, match="which"):
        tune.set_launch(ctx, 7, {"workgroups_per_cu": 2, "ray_chunk": 256, "lds_stack": 32})
    with pytest.raises(CavError, match="workgroup_size"):
        tune.set_launch(ctx, tune.STANDALONE, {"workgroups_per_cu": 2, "ray_chunk": 256, "lds_stack": 32,
                                               "workgroup_size": 96})
    with pytest.raises(CavError, match="shade_workgroup_size"):
        tune.set_launch(ctx, tune.WAVEFRONT, {"workgroups_per_cu": 2, "ray_chunk": 256, "lds_stack": 32,
                                              "shade_workgroup_size": 512})
    # 0 stands for 256
    saved = tune.get_launch(ctx, tune.STANDALONE)
    tune.set_launch(ctx, tune.STANDALONE, {"workgroups_per_cu": 2, "ray_chunk": 256, "lds_stack": 32})
    got = tune.get_launch(ctx, tune.STANDALONE)
    assert got["workgroup_size"] == 256 and got["shade_workgroup_size"] == 256
    tune.set_launch(ctx, tune.STANDALONE, saved)


@pytest.mark.parametrize("integrator", ["path", "rlpath"])
def test_render_same_under_every_workgroup_size(ctx, integrator):
    """The render loops' traversal and shading workgroup sizes and the
    cooperative passes' hand-off (1 step: every ray of this small render
    goes through them; 2^30: none) change only the speed: the film is
    bit-identical (the RL integrator in its deterministic table-update
    mode)."""
    print(f"[tune] render sizes {integrator}", flush=True)
    sc = scenes.cornell_box(seed=3, width=40, height=32)
    scene = R.Scene(ctx, sc)
    saved = tune.get_launch(ctx, tune.WAVEFRONT)
    kw = {"deterministic": True} if integrator == "rlpath" else {}
    ref = None
    try:
        for size, shade, per_cu, hand in ((256, 256, 0, 1 << 30), (64, 64, 8, 1), (128, 128, 3, 1024),
                                          (256, 64, 2, 1), (64, 128, 0, 16)):
            tune.set_launch(ctx, tune.WAVEFRONT, dict(saved, workgroup_size=size, shade_workgroup_size=shade,
                                                      workgroups_per_cu=per_cu, handoff_steps=hand))
            img, _ = R.render(ctx, scene, 2, integrator=integrator, **kw)
            if ref is None:
                ref = img.rgb.copy()
            np.testing.assert_array_equal(img.rgb, ref)
    finally:
        tune.set_launch(ctx, tune.WAVEFRONT, saved)
        scene.close()


def test_tuner_picks_caches_and_reuses(ctx, traced, tmp_path):
    bvh, rays, srays, lim = traced
    print("[tune] tune_trace", flush=True)
    dev = torch.device("cuda:0")
    r = torch.from_numpy(rays).to(dev)
    s = torch.from_numpy(srays).to(dev)
    l = torch.from_numpy(lim).to(dev)
    saved = tune.get_launch(ctx, tune.STANDALONE)
    cache = str(tmp_path / "tune.json")
    out = tune.tune_trace(ctx, bvh, r, len(rays), shadow_rays=s, limit=l, reps=1, cache_path=cache)
    assert not out["cached"] and tune.get_launch(ctx, tune.STANDALONE) == out["shape"]
    assert out["ms"] == min(out["timings"].values())
    again = tune.tune_trace(ctx, bvh, r, len(rays), shadow_rays=s, limit=l, cache_path=cache)
    assert again["cached"] and again["shape"] == out["shape"]
    # the render loops' shape, on a small render; the film is the same under any shape
    sc = scenes.cornell_box(seed=2, width=48, height=32)
    scene = R.Scene(ctx, sc)
    before, _ = R.render(ctx, scene, 2, integrator="path")
    wsaved = tune.get_launch(ctx, tune.WAVEFRONT)
    rr = tune.tune_render(ctx, scene, 2, integrator="path", reps=1, cache_path=cache)
    after, _ = R.render(ctx, scene, 2, integrator="path")
    np.testing.assert_array_equal(after.rgb, before.rgb)
    print(f"[tune] standalone {out['shape']} render {rr['shape']}", flush=True)
    tune.set_launch(ctx, tune.STANDALONE, saved)
    tune.set_launch(ctx, tune.WAVEFRONT, wsaved)
    scene.close()


def test_tune_rounds_recommends_and_caches(ctx, tmp_path):
    """tune_rounds times build + traces per candidate round count, returns
    the fastest without setting anything, and caches the recommendation;
    more rounds never raise the tree's SAH cost."""
    print("[tune] tune_rounds", flush=True)
    tris = scenes.gmm_triangles(1 << 14, seed=8)
    rays = scenes.random_rays(1 << 13, seed=9)
    cache = str(tmp_path / "tune.json")
    out = tune.tune_rounds(ctx, tris, rays, candidates=(0, 1, 2), cache_path=cache)
    assert not out["cached"] and out["rounds"] in (0, 1, 2)
    assert set(out["timings"]) == {"0", "1", "2"}
    t = out["timings"]
    assert out["ms"] == min(v["total_ms"] for v in t.values())
    assert t["2"]["sah"] <= t["1"]["sah"] <= t["0"]["sah"]
    again = tune.tune_rounds(ctx, tris, rays, candidates=(0, 1, 2), cache_path=cache)
    assert again["cached"] and again["rounds"] == out["rounds"]
