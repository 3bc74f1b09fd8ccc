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
