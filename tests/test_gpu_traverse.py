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


def _deep_chain():
    """125 triangles whose 63-bit Morton codes are 0 (65 of them) and the 60
    powers of two 2^0 .. 2^59: the Karras tree is a 60-level caterpillar
    over an index-split subtree, and every box contains the point
    (0.6, 0.5, 0.5). Box centroids sit at 0.5 + 2^(j-21) on one axis in a
    scene box of [0, 1]^3, so the reference's quantisation gives 2^j exactly."""
    def tri(c, a):
        c = np.asarray(c, np.float64)
        return np.concatenate([c + (-a, -a, -a), c + (a, a, -a), c + (0, 0, a)])
    t = []
    for axis in range(3):
        for j in range(20):
            c = [0.5, 0.5, 0.5]
            c[axis] += 2.0 ** (j - 21)
            t.append(tri(c, 0.25))
    t.append(tri([0.25, 0.25, 0.25], 0.25))  # makes the scene box [0, 1]^3
    for i in range(64):
        t.append(tri([0.5 - 0.002 * (i + 1), 0.49, 0.49], 0.25))
    return np.stack(t).astype(np.float32)


def test_stack_overflow_reports_the_reference_answers(ctx):
    """Rays from inside every box of a tree deeper than the reference's
    64-entry stack: pushes past the 64th entry are dropped, as the reference
    drops them (BVH/BVH.h:41,122), so the answers are the reference's partial
    ones, bit for bit (any-hit in the reference's whole-line order).
    cav_trace_closest / cav_trace_occluded return CAV_ERR_OVERFLOW and the
    counters hold the oracle's count of dropped pushes."""
    import ctypes

    from cavalry_amd._lib import lib
    print("[trace overflow] 125 triangles, 63-bit codes, 4096 rays", flush=True)
    tris = _deep_chain()
    codes, _, nd = oracle.lbvh_build(tris, 63)
    assert len(np.unique(codes)) == 61
    rng = np.random.default_rng(1)
    n = 4096
    o = np.float32([0.6, 0.5, 0.5]) + rng.uniform(-0.01, 0.01, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.concatenate([o, d], 1), np.float32)
    lim = np.full(n, np.inf, np.float32)
    lim[::2] = rng.uniform(0.05, 1.0, n // 2).astype(np.float32)
    ref = oracle.trace_closest(nd, tris, rays)
    assert ref["overflow"] > 0
    vref = np.empty(n, np.uint8)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    vovf = oracle.lib().ref_trace_occluded(p(nd.box), p(nd.left), p(nd.right), p(nd.prim), p(tris), None,
                                           p(rays), p(lim), None, None, ctypes.c_int64(n), p(vref))
    bvh = BVH.build(ctx, tris, optimization_rounds=0, morton_bits=63)
    prim = np.empty(n, np.int32)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    bvh.trace_counters(reset=True)
    rc = lib().cav_trace_closest(ctx.handle, bvh.handle, p(rays), n, p(prim), p(t), p(u), p(v))
    c = bvh.trace_counters(reset=True)
    # any-hit along the whole line, the reference's visited set (the limit
    # cull visits fewer boxes, so it drops other pushes, or none)
    from cavalry_amd import tune
    saved = tune.get_launch(ctx, tune.STANDALONE)
    vis = np.empty(n, np.uint8)
    try:
        tune.set_launch(ctx, tune.STANDALONE, dict(saved, any_hit_whole_line=1))
        rcv = lib().cav_trace_occluded(ctx.handle, bvh.handle, p(rays), p(lim), None, None, n, p(vis))
    finally:
        tune.set_launch(ctx, tune.STANDALONE, saved)
    cv = bvh.trace_counters(reset=True)
    print(f"[trace overflow] closest rc {rc} dropped {c['overflow_closest']} (oracle {ref['overflow']}); "
          f"any-hit rc {rcv} dropped {cv['overflow_occluded']} (oracle {vovf})", flush=True)
    assert rc == 4 and c["overflow_closest"] == ref["overflow"]
    np.testing.assert_array_equal(prim, ref["prim"])
    np.testing.assert_array_equal(bits(t), bits(ref["t"]))
    assert vovf > 0 and rcv == 4 and cv["overflow_occluded"] == vovf
    np.testing.assert_array_equal(vis, vref)
