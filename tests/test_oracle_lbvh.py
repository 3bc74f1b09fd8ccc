"""Pins for the CPU oracle's LBVH build (no GPU).

The reference ships no tests or golden vectors for this path, so the oracle
is checked against what the mathematics fixes: hand-computed Morton codes
(including the saturating cast and the clamp quirk of BVH/BVH.cu:61-62), an
independent bit-by-bit interleave, a stable argsort, the unique radix tree
over (code, index) keys built top-down, and brute-force subtree boxes.
"""
import json
import os

import numpy as np
import pytest

import oracle
from cavalry_amd import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def point_tris(points):
    """Degenerate triangles whose box (and centroid) is the given point."""
    p = np.asarray(points, np.float32)
    return np.concatenate([p, p, p], axis=1)


def interleave(x, y, z, b):
    """Bit-by-bit Morton interleave: x_i -> 3i+2, y_i -> 3i+1, z_i -> 3i."""
    c = 0
    for i in range(b):
        c |= ((x >> i) & 1) << (3 * i + 2)
        c |= ((y >> i) & 1) << (3 * i + 1)
        c |= ((z >> i) & 1) << (3 * i)
    return c


def test_morton_known_answers():
    with open(os.path.join(GOLDEN, "morton_known_answers.json")) as f:
        cases = json.load(f)
    for case in cases:
        tris = point_tris([case["point"]])
        tris = np.concatenate([tris, tris])  # n >= 1 is enough for codes
        scene = np.array(case["scene"], np.float32)
        for bits in (30, 63):
            got = int(oracle.morton(tris, bits, scene)[0])
            assert got == int(case[f"code{bits}"], 16), (case, bits, hex(got))


def test_morton_matches_bitwise_interleave():
    rng = np.random.default_rng(0)
    pts = rng.uniform(0.0, 1.0, (2000, 3)).astype(np.float32)
    scene = np.array([0, 0, 0, 1, 1, 1], np.float32)
    tris = point_tris(pts)
    for bits, b in ((30, 10), (63, 21)):
        got = oracle.morton(tris, bits, scene)
        # (c - 0.5) / 1 in float32, times 2^b, truncated; negatives saturate to 0
        c = (pts - np.float32(0.5)) / np.float32(1.0)
        q = np.where(c > 0, np.floor(c * np.float32(2.0 ** b)), 0).astype(np.int64)
        for i in range(len(pts)):
            assert int(got[i]) == interleave(int(q[i, 0]), int(q[i, 1]), int(q[i, 2]), b)


@pytest.mark.parametrize("bits", [30, 63])
def test_sort_is_stable_argsort(bits):
    tris = scenes.gmm_triangles(20000, seed=4)
    codes = oracle.morton(tris, bits)
    sc, order, _ = oracle.lbvh_build(tris, bits)
    ref = np.argsort(codes, kind="stable")
    np.testing.assert_array_equal(order, ref)
    np.testing.assert_array_equal(sc, codes[ref])


def radix_tree(sorted_codes):
    """Top-down radix tree over distinct keys (code << 32 | position): split each
    range where the highest bit differing between its ends changes. Node ids
    follow the Karras convention (internal node named by the split side it
    ends on; leaves at n-1+k)."""
    n = len(sorted_codes)
    aug = [(int(c) << 32) | i for i, c in enumerate(sorted_codes)]
    left = np.full(2 * n - 1, -1, np.int64)
    right = np.full(2 * n - 1, -1, np.int64)
    stack = [(0, 0, n - 1)]
    while stack:
        nid, lo, hi = stack.pop()
        hb = (aug[lo] ^ aug[hi]).bit_length() - 1
        a, b = lo, hi  # first index in (lo, hi] with bit hb set
        while a + 1 < b:
            mid = (a + b) // 2
            if (aug[mid] >> hb) & 1:
                b = mid
            else:
                a = mid
        g = a
        lid = n - 1 + g if g == lo else g
        rid = n - 1 + g + 1 if g + 1 == hi else g + 1
        left[nid], right[nid] = lid, rid
        if g != lo:
            stack.append((g, lo, g))
        if g + 1 != hi:
            stack.append((g + 1, g + 1, hi))
    return left, right


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("kind", ["gmm", "dup", "identical", "tiny"])
def test_karras_equals_topdown_radix_tree(bits, kind):
    if kind == "gmm":
        tris = scenes.gmm_triangles(20000, seed=8)
    elif kind == "dup":  # most centroids below the scene centre: code 0
        rng = np.random.default_rng(2)
        tris = rng.uniform(0, 0.4, (3000, 9)).astype(np.float32)
        tris[0, :3] = 1.0
    elif kind == "identical":
        tris = np.tile(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), (513, 1))
    else:
        tris = scenes.gmm_triangles(2, seed=1)
    sc, order, nd = oracle.lbvh_build(tris, bits)
    left, right = radix_tree(sc)
    n = len(tris)
    np.testing.assert_array_equal(nd.left[: n - 1], left[: n - 1])
    np.testing.assert_array_equal(nd.right[: n - 1], right[: n - 1])
    for i in range(n - 1):
        assert nd.parent[left[i]] == i and nd.parent[right[i]] == i
    assert nd.parent[0] == 0


def leaf_sets(nd):
    """Leaf node ids under every node (iterative post-order)."""
    m = len(nd.prim)
    out = [None] * m
    order = []
    stack = [0]
    while stack:
        v = stack.pop()
        order.append(v)
        if nd.prim[v] < 0:
            stack += [nd.left[v], nd.right[v]]
    for v in reversed(order):
        out[v] = [v] if nd.prim[v] >= 0 else out[nd.left[v]] + out[nd.right[v]]
    return out


def check_tree(nd, tris, exact_cost=True):
    n = nd.n
    m = 2 * n - 1
    # topology: a binary tree over all ids with every triangle in one leaf
    assert sorted(nd.prim[nd.prim >= 0].tolist()) == list(range(n))
    internal = np.nonzero(nd.prim < 0)[0]
    assert len(internal) == n - 1
    seen = np.zeros(m, int)
    for i in internal:
        seen[nd.left[i]] += 1
        seen[nd.right[i]] += 1
        assert nd.parent[nd.left[i]] == i and nd.parent[nd.right[i]] == i
    assert seen[0] == 0 and (seen[1:] == 1).all()
    # boxes: leaf box = triangle box; internal box = min/max over its leaves
    v = tris.reshape(n, 3, 3)
    tbox = np.concatenate([v.min(axis=1), v.max(axis=1)], axis=1)
    leaves = np.nonzero(nd.prim >= 0)[0]
    np.testing.assert_array_equal(nd.box[leaves], tbox[nd.prim[leaves]])
    sets = leaf_sets(nd)
    for i in internal[:: max(1, len(internal) // 2000)]:
        lb = nd.box[sets[i]]
        np.testing.assert_array_equal(nd.box[i, :3], lb[:, :3].min(axis=0))
        np.testing.assert_array_equal(nd.box[i, 3:], lb[:, 3:].max(axis=0))
    # surface area: 2(xy + yz + xz) in float32, summed left to right
    e = nd.box[:, 3:] - nd.box[:, :3]
    sa = np.float32(2) * (e[:, 0] * e[:, 1] + e[:, 1] * e[:, 2] + e[:, 0] * e[:, 2])
    np.testing.assert_array_equal(nd.sa, sa)
    # cost: closed form of the recursion, SAH = sum_leaf SA + 1.2 sum_internal SA
    closed = nd.sa[leaves].astype(np.float64).sum() + 1.2 * nd.sa[internal].astype(np.float64).sum()
    assert abs(nd.cost[0] - closed) <= 1e-5 * closed
    np.testing.assert_array_equal(nd.cost[leaves], nd.sa[leaves])


@pytest.mark.parametrize("bits", [30, 63])
@pytest.mark.parametrize("n", [2, 3, 7, 100, 5000])
def test_refit_brute_force(bits, n):
    tris = scenes.gmm_triangles(n, seed=n)
    _, _, nd = oracle.lbvh_build(tris, bits)
    check_tree(nd, tris)


def test_scene_bounds():
    tris = scenes.gmm_triangles(3000, seed=3)
    b = oracle.scene_bounds(tris)
    v = tris.reshape(-1, 3, 3)
    np.testing.assert_array_equal(b[:3], v.min(axis=(0, 1)))
    np.testing.assert_array_equal(b[3:], v.max(axis=(0, 1)))
