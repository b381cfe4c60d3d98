"""Writes bvh_fixtures.json: small PrimBounds inputs with node arrays derived by
hand from the rules of the reference builder (bvh.ts:111-296), not by running
any builder.

The rules, as bvh.ts states them:
  * buildNode(start, end) numbers nodes in DFS pre-order (bvh.ts:112); the box
    is the union of the primitives' bounds, nudged outward by |x|*1.2e-7 + 1e-9
    and stored as float32 (bvh.ts:142-147).
  * count <= 4: a leaf, skip = index + 1, leaf word start*8 + count (bvh.ts:150-153).
  * otherwise binned SAH (bvh.ts:168-281): per axis x, y, z with centroid extent
    > 1e-12, bucket = clamp(floor((c - cmin) * bins / extent), 0, bins-1); split
    s puts buckets 0..s left; cost = nL*SA(box L) + nR*SA(box R) with
    SA = 2(dx dy + dy dz + dz dx); the first strict minimum in axis-then-s
    order wins; then the two-pointer partition (bvh.ts:262-278) of order[].
  * no SAH split: medianSplit (bvh.ts:284-296): widest centroid axis (ties to
    x, then y), stable sort by centroid, mid = (start + end) >> 1.
  * interior: leaf word -1, skip = node count after both subtrees (bvh.ts:161-164).

Each fixture below lists the expected order and, per node, which primitives
its box covers, its skip and its leaf word, with the arithmetic that decides
them in the comments.  Only the float32 boxes are computed here, from those
primitive sets by the nudge rule.
"""
import json
import os

import numpy as np


def cube(c, h):
    c = np.asarray(c, np.float64)
    return c - h, c + h, c.copy()


def nudged_box(bmin, bmax, prims):
    lo = np.min(bmin[prims], axis=0)
    hi = np.max(bmax[prims], axis=0)
    out = np.zeros(8, np.float32)
    for k in range(3):
        out[k] = np.float32(lo[k] - abs(lo[k]) * 1.2e-7 - 1e-9)
        out[4 + k] = np.float32(hi[k] + abs(hi[k]) * 1.2e-7 + 1e-9)
    return out


def fixture(name, bins, prims, order, nodes, why):
    bmin = np.array([p[0] for p in prims])
    bmax = np.array([p[1] for p in prims])
    bcen = np.array([p[2] for p in prims])
    arr = []
    for cover, skip, leaf in nodes:
        nd = nudged_box(bmin, bmax, cover)
        nd[3] = skip
        nd[7] = leaf
        arr.append(nd)
    arr = np.concatenate(arr)
    return {"name": name, "bins": bins, "why": why,
            "bmin": bmin.ravel().tolist(), "bmax": bmax.ravel().tolist(),
            "bcen": bcen.ravel().tolist(), "order": order, "nodeCount": len(nodes),
            "nodes_u32": arr.view(np.uint32).tolist()}


fixtures = []

# A. Five unit cubes on the x axis at x = 0..4 (y = z = 0).  y and z extents
# are 0, so only x is binned: toBucket = 12/4 = 3, buckets 0, 3, 6, 9, 11.
# With SA(cube run of length L) = 2(L*1 + 1*1 + 1*L) = 4L + 2:
#   s = 0..2  : L = {0}        cost 1*6  + 4*18 = 78   (first: best 78)
#   s = 3..5  : L = {0,1}      cost 2*10 + 3*14 = 62   (s = 3 strictly better)
#   s = 6..8  : L = {0,1,2}    cost 3*14 + 2*10 = 62   (equal: not taken)
#   s = 9,10  : L = {0..3}     cost 4*18 + 1*6  = 78
# so axis x, split 3: prims 0,1 go left.  Two-pointer on [0,1,2,3,4]: lo
# passes 0,1; at lo=2 prim 2 is bad -> swap with hi=4 -> [0,1,4,3,2], hi=3;
# prim 4 bad -> swap with 3 -> [0,1,3,4,2], hi=2; prim 3 bad -> swap with
# itself, hi=1; stop, mid = 2.  Children [0,2) and [2,5) are leaves.
fixtures.append(fixture(
    "line5_equal_costs", 12, [cube((x, 0, 0), 0.5) for x in range(5)],
    [0, 1, 3, 4, 2],
    [([0, 1, 2, 3, 4], 3, -1),
     ([0, 1], 2, 0 * 8 + 2),
     ([3, 4, 2], 3, 2 * 8 + 3)],
    "equal SAH costs at different splits of one axis: the first strict minimum wins"))

# B. Five unit cubes at (x, z) = (0,0), (0,4), (4,0), (4,4), (2,2), y = 0 --
# symmetric under swapping x and z.  Per axis x (and likewise z): toBucket =
# 12/4 = 3, buckets 0, 0, 11, 11, 6.
#   s = 0..5  : L = {P0,P1} (dx 1, dz 5: SA 22), R = {P2,P3,P4} (dx 3, dz 5:
#               SA 46): cost 2*22 + 3*46 = 182
#   s = 6..10 : L = {P0,P1,P4} (SA 46), R = {P2,P3} (SA 22): cost 182
# z gives the same costs with P1 and P2 exchanged; y is skipped (extent 0).
# The first strict minimum is x, s = 0: P0, P1 left.  Two-pointer on
# [0,1,2,3,4]: lo passes 0,1; P2 bad -> swap with 4 -> [0,1,4,3,2]; P4 bad ->
# swap with 3 -> [0,1,3,4,2]; P3 bad -> itself; mid = 2.
fixtures.append(fixture(
    "axis_tie5", 12, [cube(c, 0.5) for c in ((0, 0, 0), (0, 0, 4), (4, 0, 0), (4, 0, 4),
                                             (2, 0, 2))],
    [0, 1, 3, 4, 2],
    [([0, 1, 2, 3, 4], 3, -1),
     ([0, 1], 2, 2),
     ([3, 4, 2], 3, 2 * 8 + 3)],
    "equal costs across axes and within an axis: x before z, first split before later"))

# C. Three far cubes at x = -10, -9, -8 (indices 1, 3, 6) and six cubes of a
# cluster at x = 10 + ox, z = 10 + oz with offsets below 1e-12 (indices 0, 2,
# 4, 5, 7, 8); every centroid has y = 0 and the far ones z = 10.
# Root (9 prims): y and z extents are 0 and 3e-13 (skipped); x: toBucket =
# 12/(20 + ox_max) ~ 0.6, far buckets 0, 0, 1, cluster 11.
#   s = 0      : L = {-10,-9}: cost 2*SA + 7*SA(rest) ~ 2*10 + 7*164
#   s = 1..10  : L = far three (SA 14), R = cluster (a 2-unit cube, SA 24):
#                cost 3*14 + 6*24 = 186, the minimum, first at s = 1.
# Two-pointer on [0..8] (bad = cluster): prim 0 bad -> swap with 8 ->
# [8,1,2,3,4,5,6,7,0]; 8 bad -> swap with 7 -> [7,1,...,8,0]; 7 bad -> swap
# with 6 -> [6,1,2,3,4,5,7,8,0]; 6 good; 1 good; 2 bad -> swap with 5 ->
# [6,1,5,3,4,2,...]; 5 bad -> swap with 4 -> [6,1,4,3,5,2,...]; 4 bad -> swap
# with 3 -> [6,1,3,4,5,2,7,8,0]; 3 good; mid = 3.
# Left [0,3) = {6,1,3}: a leaf.  Right [3,9) = [4,5,2,7,8,0]: every extent is
# <= 1e-12, so medianSplit: ex = ez (the same offsets, 0..3e-13, on both axes
# around 10) > ey = 0, so x (ties go to x).  Stable sort by x offset:
# 4:2, 5:0, 2:0, 7:1, 8:3, 0:3 (units of 1e-13) -> [5,2,7,4,8,0]; mid =
# (3+9)>>1 = 6: leaves [5,2,7] and [4,8,0].
ox = {0: 3e-13, 2: 0.0, 4: 2e-13, 5: 0.0, 7: 1e-13, 8: 3e-13}
oz = {0: 0.0, 2: 3e-13, 4: 1e-13, 5: 3e-13, 7: 2e-13, 8: 0.0}
half = {0: 1.0, 2: 0.8, 4: 0.6, 5: 0.4, 7: 0.3, 8: 0.2}
far = {1: -10.0, 3: -9.0, 6: -8.0}
prims_c = []
for i in range(9):
    if i in far:
        prims_c.append(cube((far[i], 0.0, 10.0), 0.5))
    else:
        prims_c.append(cube((10.0 + ox[i], 0.0, 10.0 + oz[i]), half[i]))
fixtures.append(fixture(
    "sah_then_median9", 12, prims_c,
    [6, 1, 3, 5, 2, 7, 4, 8, 0],
    [(list(range(9)), 5, -1),
     ([6, 1, 3], 2, 0 * 8 + 3),
     ([5, 2, 7, 4, 8, 0], 5, -1),
     ([5, 2, 7], 4, 3 * 8 + 3),
     ([4, 8, 0], 5, 6 * 8 + 3)],
    "SAH partition then a median split: widest-axis tie to x, stable order of equal centroids"))

# D. Six cubes of one centroid (every extent 0): median on x, the stable sort
# keeps [0..5], mid = 3: leaves [0,3) and [3,6) -- bvh.test.ts:153-168's case
# at a size where the split is visible.
fixtures.append(fixture(
    "coincident6", 12, [cube((1, 2, 3), h) for h in (0.5, 0.1, 0.9, 0.3, 0.7, 0.2)],
    [0, 1, 2, 3, 4, 5],
    [([0, 1, 2, 3, 4, 5], 3, -1),
     ([0, 1, 2], 2, 3),
     ([3, 4, 5], 3, 3 * 8 + 3)],
    "all centroids equal: median split keeps the input order"))

here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "bvh_fixtures.json"), "w") as f:
    json.dump(fixtures, f, indent=1)
print(f"wrote {len(fixtures)} fixtures")
