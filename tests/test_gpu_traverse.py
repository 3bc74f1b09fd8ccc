"""Closest-hit and any-hit traversal on the GPU vs the CPU oracle: hit ids
exact, t/u/v bit for bit, visibility exact."""
import numpy as np
import pytest

import oracle
from cavalry_amd import BVH, scenes
from helpers import bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rounds", [0, 3])
@pytest.mark.parametrize("n,nrays,edge", [(2, 1000, 0.3), (64, 5000, 0.1), (10000, 50000, 1e-3),
                                          (10000, 50000, 0.02), (200000, 100000, 1e-3),
                                          (200000, 100000, 0.01)])
def test_closest_and_occluded(ctx, n, nrays, edge, rounds):
    print(f"[trace] n={n} rays={nrays} edge={edge} rounds={rounds}", flush=True)
    tris = scenes.gmm_triangles(n, seed=n + 1, edge=edge)
    _, _, nd = oracle.lbvh_build(tris, 30)
    if rounds:
        oracle.treelet_optimize(nd, rounds)
    rays = scenes.random_rays(nrays, seed=n + 2)
    ref = oracle.trace_closest(nd, tris, rays)
    bvh = BVH.build(ctx, tris, optimization_rounds=rounds)
    got = bvh.intersect(rays)
    np.testing.assert_array_equal(got["prim"], ref["prim"])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(bits(got[k]), bits(ref[k]), err_msg=k)
    assert (ref["prim"] >= 0).any() or n < 100

    srays = scenes.random_rays(nrays, seed=n + 3, stream=2)
    lim = scenes.shadow_limits(nrays, seed=n + 4)
    lim[::7] = np.inf
    vref = oracle.trace_occluded(nd, tris, srays, lim)
    vgot = bvh.test_visibility(srays, lim)
    np.testing.assert_array_equal(vgot, vref)


@pytest.mark.parametrize("step", [0, 1, 2])
def test_occluded_mesh_exclusion(ctx, step):
    """Both any-hit kernels (one node or a sibling pair per step; 0: one
    node) skip the source and target meshes as the reference does."""
    from cavalry_amd import tune
    n = 5000
    tris = scenes.gmm_triangles(n, seed=9)
    mesh = (np.arange(n) % 13).astype(np.int32)
    _, _, nd = oracle.lbvh_build(tris, 30)
    rays = scenes.random_rays(20000, seed=10)
    lim = scenes.shadow_limits(20000, seed=11)
    rng = np.random.default_rng(12)
    src = rng.integers(-1, 13, 20000).astype(np.int32)
    dst = rng.integers(-1, 13, 20000).astype(np.int32)
    vref = oracle.trace_occluded(nd, tris, rays, lim, mesh, src, dst)
    bvh = BVH.build(ctx, tris, tri_mesh=mesh)
    saved = tune.get_launch(ctx, tune.STANDALONE)
    try:
        tune.set_launch(ctx, tune.STANDALONE, dict(saved, any_hit_step=step))
        vgot = bvh.test_visibility(rays, lim, src, dst)
    finally:
        tune.set_launch(ctx, tune.STANDALONE, saved)
    np.testing.assert_array_equal(vgot, vref)


@pytest.mark.parametrize("step", [1, 2])
@pytest.mark.parametrize("n,nrays,edge,rounds", [(10000, 50000, 1e-3, 0), (200000, 100000, 0.01, 3),
                                                 (4000, 30000, 0.3, 1)])
def test_every_ray_through_the_cooperative_passes(ctx, n, nrays, edge, rounds, step):
    """With a hand-off after one step, the rays still in flight once the
    rays are all handed out -- here a large share: the batch is smaller than
    the persistent grid -- are traced again by a whole wave (closest-hit) or
    finished by one (any-hit, spilled by either any-hit kernel). Hits and visibility stay the oracle's,
    bit for bit; the rays the closest-hit pass sends back to one lane (two
    hits within a hair, a deep stack) are traced in the reference's order."""
    from cavalry_amd import tune
    print(f"[trace coop] n={n} rays={nrays} edge={edge} rounds={rounds} any-hit step={step}", flush=True)
    tris = scenes.gmm_triangles(n, seed=n + 11, edge=edge)
    _, _, nd = oracle.lbvh_build(tris, 30)
    if rounds:
        oracle.treelet_optimize(nd, rounds)
    rays = scenes.random_rays(nrays, seed=n + 12)
    srays = scenes.random_rays(nrays, seed=n + 13, stream=2)
    lim = scenes.shadow_limits(nrays, seed=n + 14)
    lim[::5] = np.inf
    ref = oracle.trace_closest(nd, tris, rays)
    vref = oracle.trace_occluded(nd, tris, srays, lim)
    bvh = BVH.build(ctx, tris, optimization_rounds=rounds)
    saved = tune.get_launch(ctx, tune.STANDALONE)
    try:
        tune.set_launch(ctx, tune.STANDALONE, dict(saved, handoff_steps=1, any_hit_step=step))
        got = bvh.intersect(rays)
        f = bvh.trace_fetches()
        vgot = bvh.test_visibility(srays, lim)
        fv = bvh.trace_fetches()
    finally:
        tune.set_launch(ctx, tune.STANDALONE, saved)
    print(f"[trace coop] handed over {f['long_closest_rays']} closest "
          f"(retraced {f['long_closest_retraced']}), {fv['spilled_rays']} any-hit", flush=True)
    np.testing.assert_array_equal(got["prim"], ref["prim"])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(bits(got[k]), bits(ref[k]), err_msg=k)
    np.testing.assert_array_equal(vgot, vref)
    # (a lane asks whether the rays are all handed out at its ray's first
    # step, then every 256 steps: the rays taken after the last wave drew
    # its chunk go over)
    assert f["long_closest_rays"] > nrays // 8 and fv["spilled_rays"] > nrays // 8
    assert f["long_closest_retraced"] < f["long_closest_rays"]


def test_cooperative_passes_match_one_lane_per_ray(ctx):
    """At a size where long rays are the tail (2^18 triangles, 2^18 rays):
    the default hand-off and no hand-off at all give the same hits and
    visibility."""
    from cavalry_amd import tune
    print("[trace coop] default vs none at 2^18", flush=True)
    tris = scenes.gmm_triangles(1 << 18, seed=21)
    rays = scenes.random_rays(1 << 18, seed=22)
    srays = scenes.random_rays(1 << 18, seed=23, stream=2)
    lim = scenes.shadow_limits(1 << 18, seed=24)
    bvh = BVH.build(ctx, tris, optimization_rounds=3)
    saved = tune.get_launch(ctx, tune.STANDALONE)
    try:
        tune.set_launch(ctx, tune.STANDALONE, dict(saved, handoff_steps=1 << 30))
        a, va = bvh.intersect(rays), bvh.test_visibility(srays, lim)
        tune.set_launch(ctx, tune.STANDALONE, dict(saved, handoff_steps=64))
        b, vb = bvh.intersect(rays), bvh.test_visibility(srays, lim)
        f = bvh.trace_fetches()
    finally:
        tune.set_launch(ctx, tune.STANDALONE, saved)
    print(f"[trace coop] {f}", flush=True)
    assert f["long_closest_rays"] > 0
    np.testing.assert_array_equal(a["prim"], b["prim"])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(bits(a[k]), bits(b[k]), err_msg=k)
    np.testing.assert_array_equal(va, vb)


@pytest.mark.parametrize("step", [1, 2])
@pytest.mark.parametrize("handoff", [1024, 1])
def test_limit_cull_keeps_visibility(ctx, step, handoff):
    """Any-hit rays skip boxes that start past their limit (the default) or
    visit every box their line crosses, as the reference does: the same
    visibility as the oracle's either way, on limits from negative and zero
    to infinite, with fewer records fetched when the limits are finite."""
    from cavalry_amd import tune
    print(f"[trace cull] any-hit step={step} handoff={handoff}", flush=True)
    n, nrays = 200000, 131072
    tris = scenes.gmm_triangles(n, seed=31, edge=2e-3)
    _, _, nd = oracle.lbvh_build(tris, 30)
    oracle.treelet_optimize(nd, 1)
    srays = scenes.random_rays(nrays, seed=32, stream=2)
    lim = scenes.shadow_limits(nrays, seed=33)
    lim[::11] = np.inf
    lim[1::97] = 0.0
    lim[2::97] = -1.0
    lim[3::97] = 1e-6
    lim[4::97] = 1e30
    vref = oracle.trace_occluded(nd, tris, srays, lim)
    bvh = BVH.build(ctx, tris, optimization_rounds=1)
    saved = tune.get_launch(ctx, tune.STANDALONE)
    out = {}
    try:
        for whole in (0, 1):
            tune.set_launch(ctx, tune.STANDALONE, dict(saved, handoff_steps=handoff, any_hit_step=step,
                                                       any_hit_whole_line=whole))
            bvh.trace_fetches(reset=True)
            out[whole] = (bvh.test_visibility(srays, lim), bvh.trace_fetches(reset=True)["occluded"])
    finally:
        tune.set_launch(ctx, tune.STANDALONE, saved)
    print(f"[trace cull] records culled {out[0][1]} whole line {out[1][1]}", flush=True)
    np.testing.assert_array_equal(out[1][0], vref)
    np.testing.assert_array_equal(out[0][0], vref)
    assert out[0][1] < out[1][1]
