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


def test_occluded_mesh_exclusion(ctx):
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
    vgot = bvh.test_visibility(rays, lim, src, dst)
    np.testing.assert_array_equal(vgot, vref)
