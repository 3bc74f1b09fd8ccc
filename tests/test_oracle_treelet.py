"""Pins for the CPU oracle's treelet restructuring (no GPU).

* the hard-coded partition schedule covers every subset of 2..5 of the 7
  leaves exactly once, each after all of its sub-subsets (parsed from the
  oracle and the HIP source, which must hold the same table);
* on 7-leaf trees one round reaches the optimum SAH found by brute force over
  all (2*7-3)!! = 10395 binary trees;
* after rounds on larger trees the result is still a valid BVH over the same
  triangles, boxes/areas are exact, and the root SAH never goes up.
"""
import itertools
import os
import re

import numpy as np
import pytest

import oracle
from cavalry_amd import scenes
from test_oracle_lbvh import check_tree

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def schedule(path):
    text = open(path).read()
    body = re.search(r"kSchedule\[160\] = \{(.*?)\};", text, re.S).group(1)
    vals = [int(t, 0) for t in re.findall(r"0b[01]+|\b\d+\b", body)]
    assert len(vals) == 160
    return vals


def test_schedule_tables_identical_and_valid():
    a = schedule(os.path.join(ROOT, "oracle", "ref_treelet.cpp"))
    b = schedule(os.path.join(ROOT, "cavalry_amd", "csrc", "treelet.hip"))
    assert a == b
    rounds = [[s for s in a[r * 32:(r + 1) * 32] if s] for r in range(5)]
    flat = [s for r in rounds for s in r]
    want = [s for s in range(1, 128) if 2 <= bin(s).count("1") <= 5]
    assert sorted(flat) == want  # each subset of size 2..5 exactly once
    done = {1 << i for i in range(7)}
    for r in rounds:
        for s in r:
            sub = s & (s - 1)
            while sub:  # every proper non-empty subset must be finished
                if sub != s:
                    assert sub in done, (bin(s), bin(sub))
                sub = (sub - 1) & s
        done |= set(r)


def all_trees(leaves):
    """Every full binary tree over the given leaf labels (nested tuples)."""
    if len(leaves) == 1:
        yield leaves[0]
        return
    first, rest = leaves[0], leaves[1:]
    for k in range(0, len(rest)):
        for left_rest in itertools.combinations(rest, k):
            left = (first,) + left_rest
            right = tuple(x for x in rest if x not in left_rest)
            for lt in all_trees(list(left)):
                for rt in all_trees(list(right)):
                    yield (lt, rt)


def sah(tree, boxes):
    """(cost, box) with cost(leaf) = SA, cost(node) = 1.2 SA + cost(l) + cost(r)."""
    if not isinstance(tree, tuple):
        b = boxes[tree].astype(np.float64)
        e = b[3:] - b[:3]
        return 2 * (e[0] * e[1] + e[1] * e[2] + e[0] * e[2]), b
    cl, bl = sah(tree[0], boxes)
    cr, br = sah(tree[1], boxes)
    b = np.concatenate([np.minimum(bl[:3], br[:3]), np.maximum(bl[3:], br[3:])])
    e = b[3:] - b[:3]
    return 1.2 * 2 * (e[0] * e[1] + e[1] * e[2] + e[0] * e[2]) + cl + cr, b


@pytest.mark.parametrize("seed", range(6))
def test_seven_leaves_reach_brute_force_optimum(seed):
    rng = np.random.default_rng(seed)
    tris = (rng.uniform(0, 1, (7, 1, 3)) + rng.normal(0, 0.05, (7, 3, 3))).reshape(7, 9)
    tris = tris.astype(np.float32)
    _, _, nd = oracle.lbvh_build(tris, 30)
    v = tris.reshape(7, 3, 3)
    boxes = np.concatenate([v.min(axis=1), v.max(axis=1)], axis=1)
    trees = list(all_trees(list(range(7))))
    assert len(trees) == 10395
    best = min(sah(t, boxes)[0] for t in trees)
    before = float(nd.cost[0])
    assert oracle.treelet_optimize(nd, 1) == 1
    check_tree(nd, tris)
    assert abs(float(nd.cost[0]) - best) <= 1e-5 * best, (nd.cost[0], best, before)


@pytest.mark.parametrize("n,rounds", [(8, 2), (100, 3), (3000, 3), (20000, 2)])
def test_rounds_keep_a_valid_tree_and_lower_sah(n, rounds):
    tris = scenes.gmm_triangles(n, seed=n + 7)
    _, _, nd = oracle.lbvh_build(tris, 30)
    costs = [float(nd.cost[0])]
    for _ in range(rounds):
        oracle.treelet_optimize(nd, 1)
        check_tree(nd, tris)
        costs.append(float(nd.cost[0]))
    for a, b in zip(costs, costs[1:]):
        assert b <= a * (1 + 1e-6), costs
    assert costs[-1] < costs[0]


def test_fewer_than_seven_leaves_is_a_no_op():
    tris = scenes.gmm_triangles(6, seed=2)
    _, _, nd = oracle.lbvh_build(tris, 30)
    before = nd.copy()
    assert oracle.treelet_optimize(nd, 3) == 0
    for k in ("box", "sa", "cost", "left", "right", "parent", "prim"):
        np.testing.assert_array_equal(getattr(nd, k), getattr(before, k))
