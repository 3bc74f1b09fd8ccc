"""cav_trace_closest_noalloc against cav_trace_closest: the same hits, bit
for bit, on two calls in a row (the second with fewer rays, so stale staging
from the first must not leak into it), and the size limit refused."""
import ctypes

import numpy as np
import pytest

from cavalry_amd import BVH, scenes
from cavalry_amd._lib import CavError, check, lib

pytestmark = pytest.mark.gpu

MAX_RAYS = 262144  # CAV_NOALLOC_MAX_RAYS in include/cavalry.h


def trace_noalloc(ctx, bvh, rays):
    rays = np.ascontiguousarray(rays, np.float32)
    r = len(rays)
    out = {k: np.full(r, np.nan, np.float32) for k in ("t", "u", "v")}
    out["prim"] = np.full(r, -7, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    check(ctx.handle, lib().cav_trace_closest_noalloc(ctx.handle, bvh.handle, p(rays), r, p(out["prim"]),
                                                      p(out["t"]), p(out["u"]), p(out["v"])))
    return out


def test_noalloc_matches_trace_closest(ctx):
    tris = scenes.gmm_triangles(60000, seed=41, edge=0.01)
    bvh = BVH.build(ctx, tris, optimization_rounds=1, morton_bits=30)
    for nr, seed in ((MAX_RAYS, 42), (50001, 43)):
        print(f"[noalloc] {nr} rays", flush=True)
        rays = scenes.random_rays(nr, seed=seed, lo=0.0, hi=1.0)
        want = bvh.intersect(rays)  # cav_trace_closest: per-call device staging
        got = trace_noalloc(ctx, bvh, rays)
        assert (want["prim"] >= 0).sum() > nr // 20
        np.testing.assert_array_equal(got["prim"], want["prim"])
        for k in ("t", "u", "v"):
            np.testing.assert_array_equal(got[k].view(np.uint32), want[k].view(np.uint32), err_msg=k)
    too_many = scenes.random_rays(MAX_RAYS + 1, seed=44, lo=0.0, hi=1.0)
    with pytest.raises(CavError, match="at most 262144 rays"):
        trace_noalloc(ctx, bvh, too_many)
    bvh.close()
