"""LBVH build on the GPU vs the CPU oracle: Morton codes, sorted order, tree
topology and float boxes / areas / costs must match bit for bit."""
import numpy as np
import pytest

import oracle
from cavalry_amd import BVH, scenes
from helpers import assert_nodes_equal

pytestmark = pytest.mark.gpu


def _check(ctx, tris, bits, rounds=0, scene6=None):
    print(f"[lbvh] n={len(tris)} bits={bits} rounds={rounds}", flush=True)
    codes, order, nd = oracle.lbvh_build(tris, bits, scene6)
    if rounds:
        oracle.treelet_optimize(nd, rounds)
    bvh = BVH.build(ctx, tris, scene_bounds=scene6, optimization_rounds=rounds, morton_bits=bits)
    ex = bvh.export()
    np.testing.assert_array_equal(ex["codes"], codes)
    np.testing.assert_array_equal(ex["order"], order)
    assert_nodes_equal(ex, nd, f"n={len(tris)} bits={bits} rounds={rounds}: ")
    return bvh, nd


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("n", [2, 3, 5, 7, 8, 100, 1000, 4095, 4096, 4097, 65536, 250000])
def test_lbvh_bit_exact(ctx, n, bits):
    _check(ctx, scenes.gmm_triangles(n, seed=n), bits)


@pytest.mark.parametrize("bits", [30, 63])
def test_lbvh_duplicate_keys(ctx, bits):
    # every centroid below the scene centre: all Morton codes are 0 (the
    # saturating cast), so the tree comes from the index tie-break alone
    rng = np.random.default_rng(5)
    n = 5000
    tris = (rng.uniform(0.0, 0.4, (n, 9))).astype(np.float32)
    tris[0, :3] = 1.0  # stretch the scene box so everything else is below centre
    _check(ctx, tris, bits)
    # identical triangles
    t2 = np.tile(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), (777, 1))
    _check(ctx, t2, bits)


@pytest.mark.parametrize("bits", [30, 63])
def test_lbvh_signed_zero_and_given_bounds(ctx, bits):
    # coordinates of +0 and -0 in every order exercise the min/max semantics
    rng = np.random.default_rng(7)
    n = 3000
    tris = rng.choice(np.array([-0.0, 0.0, 0.25, -0.5, 1.0], np.float32), size=(n, 9))
    tris = tris.astype(np.float32)
    _check(ctx, tris, bits)
    # caller-supplied scene bounds (the reference passes sceneBounds in)
    tris2 = scenes.gmm_triangles(2000, seed=11)
    _check(ctx, tris2, bits, scene6=np.array([-1, -1, -1, 2, 2, 2], np.float32))


@pytest.mark.parametrize("rounds", [1, 3])
@pytest.mark.parametrize("n", [7, 8, 50, 1000, 20000, 200000])
def test_treelet_bit_exact(ctx, n, rounds):
    _check(ctx, scenes.gmm_triangles(n, seed=100 + n), 30, rounds=rounds)


def test_treelet_63bit(ctx):
    _check(ctx, scenes.gmm_triangles(30000, seed=3), 63, rounds=2)


def test_build_errors(ctx):
    from cavalry_amd._lib import CavError
    with pytest.raises(CavError):
        BVH.build(ctx, np.zeros((1, 9), np.float32))
    with pytest.raises(CavError):
        BVH.build(ctx, np.zeros((4, 9), np.float32), morton_bits=31)


def test_treelet_round_hand_derived(ctx):
    # the 7-leaf case worked by hand in test_treelet_golden.py
    from test_treelet_golden import G, triangles
    bvh = BVH.build(ctx, triangles(), optimization_rounds=1, morton_bits=30)
    ex = bvh.export()
    want = G["after_one_round"]
    for k in ("left", "right", "parent", "prim"):
        np.testing.assert_array_equal(ex[k], np.asarray(want[k]), err_msg=k)
    np.testing.assert_array_equal(ex["box"][:, [0, 3]], np.asarray(want["x_extent"], np.float32))
    bvh.close()


@pytest.mark.parametrize("bits", [30, 63])
def test_rebuild_in_place_bit_exact(ctx, bits):
    # one BVH object rebuilt from other triangles of the same count: the
    # split gates the first build crossed must be clear again for the second
    n = 300000
    print(f"[lbvh] rebuild n={n} bits={bits}", flush=True)
    a = scenes.gmm_triangles(n, seed=31)
    b = scenes.gmm_triangles(n, seed=32)
    import torch
    bvh = BVH.build(ctx, a, optimization_rounds=0, morton_bits=bits)
    for tris in (b, a):
        dev = torch.from_numpy(tris).to("cuda:0")  # rebuild reads device triangles
        torch.cuda.synchronize()
        bvh.rebuild(dev, optimization_rounds=0, morton_bits=bits)
        codes, order, nd = oracle.lbvh_build(tris, bits)
        ex = bvh.export()
        np.testing.assert_array_equal(ex["codes"], codes)
        np.testing.assert_array_equal(ex["order"], order)
        assert_nodes_equal(ex, nd, f"rebuild n={n} bits={bits}: ")
    bvh.close()
