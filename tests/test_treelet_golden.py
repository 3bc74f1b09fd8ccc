"""A treelet round worked by hand (BVH/Optimization.cu:49-462).

Seven unit cubes along x, as triangles v0 = (x,0,0), v1 = (x+1,1,0),
v2 = (x,0,1), at x = 0, 10, 1.5, 11.5, 40, 41.5, 44 (triangles 0..6). Every
box has y = z = [0, 1], so a box of x-length L has area 2(2L + 1) = 4L + 2.

Morton codes (BVH/BVH.cu:79-92). Scene x = [0, 45], centre 22.5, extent 45;
y and z centroids sit at their scene centres and code 0. Centroids below the
centre saturate to 0 on that axis, so triangles 0..3 all get code 0 and keep
their index order in the stable sort. Triangles 4..6 get x codes
floor(1024 (c - 22.5) / 45) = 409, 443, 500 (0b0110011001, 0b0110111011,
0b0111110100), and the sorted order is the identity.

Karras tree (BVH/BVH.cu:122-215). Adjacent prefix lengths: delta(0..2) =
30 + clz(j ^ (j+1)) = 61, 60, 61 (equal codes); delta(3) = 3 (x bit 8 is
code bit 26); delta(4) = 12 (409 and 443 first differ at x bit 5, code bit
17); delta(5) = 9 (x bit 6, code bit 20). Root [0,6] splits at 3; [0,3] (a
left child: node 3) at 1 into node 1 = [0,1] and node 2 = [2,3]; [4,6] (a
right child: node 4) at 5 into node 5 = [4,5] (a left child) and leaf 6
(id 12); node 5 holds leaves 4, 5 (ids 10, 11). Leaves 0..6 are ids 6..12.

Treelet formation at the root (the only node with 7 leaves below it; every
smaller treelet gives up). Areas: root 182, node 3 = [0,12.5] 52, node 1 =
[0,11] 46, node 2 = [1.5,12.5] 46, node 4 = [40,45] 22, node 5 = [40,42.5]
12. Lane 0 holds the root; each step expands the largest unexpanded internal
node (the lowest lane on ties), its children going to the next two lanes:
root -> lanes 1, 2 = nodes 3, 4; node 3 -> lanes 3, 4 = nodes 1, 2; node 1
(46, before node 2's equal 46) -> lanes 5, 6 = ids 6, 7; node 2 -> lanes 7, 8
= ids 8, 9; node 4 -> lanes 9, 10 = node 5, id 12; node 5 -> lanes 11, 12 =
ids 10, 11. The treelet leaves in lane order, bits 0..6 of a subset, are
ids 6, 7, 8, 9, 12, 10, 11: x = 0, 10, 1.5, 11.5, 44, 40, 41.5.

Optimal partitions (leaf cost = area, internal = 1.2 area + children). For
subsets of 2..5 leaves the enumerated side p never holds the subset's lowest
bit and becomes the left child. {0,1,2,3}: best p = {1,3} (x 10, 11.5) against
{0,2} (x 0, 1.5), 26.4 + 26.4; the Karras pairs would cost 67.2 + 67.2.
{4,5,6} (x 44, 40, 41.5): best p = {5,6} = [40,42.5] (26.4) with {4} (6);
{1,3}: p = {3}; {0,2}: p = {2}; {5,6}: p = {6}. For the full set p never
holds bit 6: p = {0,1,2,3} (cost 1.2 52 + 52.8 = 115.2) against {4,5,6}
(1.2 22 + 32.4 = 58.8).

Reconstruction: lane 0 (root) keeps all 7; each step expands the first lane
(in lane order) with an unexpanded subset of 2 or more, its left and right
parts going to the next two lanes, whose node ids the formation fixed:
127 -> lanes 1 ({0,1,2,3}: node 3), 2 ({4,5,6}: node 4); {0,1,2,3} -> lanes
3 ({1,3}: node 1), 4 ({0,2}: node 2); {4,5,6} -> lanes 5 ({5,6}: id 6), 6
({4}: id 7); {1,3} -> lanes 7 ({3}: id 8), 8 ({1}: id 9); {0,2} -> lanes 9
({2}: id 5), 10 ({0}: id 12); {5,6} -> lanes 11 ({6}: id 10), 12 ({5}:
id 11). A lane left with one bit takes that treelet leaf's triangle:
id 7 <- triangle 6, id 8 <- 3, id 9 <- 1, id 5 <- 2, id 12 <- 0, id 10 <- 5,
id 11 <- 4. Root cost: 1.2 182 + 115.2 + 58.8 = 392.4.
"""
import json
import os

import numpy as np

import oracle

G = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "treelet_7leaf.json")))


def triangles():
    x = np.asarray(G["x"], np.float32)
    t = np.zeros((7, 9), np.float32)
    t[:, 0] = x
    t[:, 3] = x + 1
    t[:, 4] = 1
    t[:, 6] = x
    t[:, 8] = 1
    return t


def check(ex, want, what):
    for k in ("left", "right", "parent", "prim"):
        np.testing.assert_array_equal(np.asarray(ex[k]), np.asarray(want[k]), err_msg=f"{what}: {k}")


def test_oracle_karras_tree_is_the_hand_derived_one():
    codes, order, nd = oracle.lbvh_build(triangles(), 30)
    np.testing.assert_array_equal(order, np.arange(7))
    assert list(codes[:4]) == [0, 0, 0, 0]
    check({"left": nd.left, "right": nd.right, "parent": nd.parent, "prim": nd.prim}, G["before"], "Karras")


def test_oracle_treelet_round_is_the_hand_derived_one():
    _, _, nd = oracle.lbvh_build(triangles(), 30)
    oracle.treelet_optimize(nd, 1)
    want = G["after_one_round"]
    check({"left": nd.left, "right": nd.right, "parent": nd.parent, "prim": nd.prim}, want, "one round")
    np.testing.assert_array_equal(nd.box[:, [0, 3]], np.asarray(want["x_extent"], np.float32))
    np.testing.assert_allclose(nd.cost[0], want["root_cost"], rtol=1e-6)
