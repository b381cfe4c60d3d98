"""bvh.test.ts restated against the oracle (the walks are pinned by brute force,
the build by validateBvh's invariants, the partition by its closed form)."""
import math

import numpy as np
import pytest

import scenes


@pytest.fixture(scope="module")
def ps(O):
    rng = O.Rng(1234)
    return O.pack_scene(scenes.random_scene(rng, 1000, 2000))


def test_pack_counts(ps):  # bvh.test.ts:50-54
    assert ps.primCount == 3000
    assert ps.nodeCount > 1
    assert len(ps.nodes) == ps.nodeCount * 8


def test_leaf_sizes(ps):  # bvh.test.ts:56-67
    leaf = ps.nodes[7::8]
    lv = leaf[leaf >= 0].astype(np.int64)
    cnt = lv - (lv // 8) * 8
    assert (cnt <= 4).all() and (cnt > 0).all()
    assert len(lv) >= math.ceil(3000 / 4)


def test_validate_and_skip_links(O, ps):  # bvh.test.ts:69-80
    rep = O.validate_bvh(ps.nodes, ps.nodeCount, ps.order, ps.primCount)
    assert rep["errors"] == 0 and rep["maxLeafSize"] <= 4
    skip = ps.nodes[3::8]
    idx = np.arange(ps.nodeCount)
    assert (skip > idx).all() and (skip <= ps.nodeCount).all()


@pytest.mark.parametrize("walk", [0, 1], ids=["ordered", "stackless"])
def test_walk_matches_brute_force(O, ps, walk):  # bvh.test.ts:87-112
    rays = scenes.ray_set_bvh_test(O.Rng(99), 1000)
    a = O.closest_batch(ps, rays, walk=walk)
    b = O.closest_batch(ps, rays, walk=O.WALK_BRUTE)
    hit = a["prim"] >= 0
    assert (hit == (b["prim"] >= 0)).all()
    assert hit.sum() > 500
    assert (a["prim"] == b["prim"]).all()
    assert np.all(np.abs(a["t"][hit] - b["t"][hit]) < 1e-6)
    # u/v are only meaningful for triangle hits: on a sphere hit both walks keep
    # whatever an earlier triangle test left (intersect.ts:173-176, 332-335), and
    # the ordered walk tests prims in a different order than the brute-force scan
    tri = hit & (ps.prims[np.maximum(a["prim"], 0) * 16 + 3] == 1)
    assert np.all(np.abs(a["u"][tri] - b["u"][tri]) < 1e-9)
    assert np.all(np.abs(a["v"][tri] - b["v"][tri]) < 1e-9)


def _rays_finite(O, seed, n, lo, hi, tlo, thi):
    rng = O.Rng(seed)
    d = rng.take(7 * n).reshape(n, 7)
    o = lo + (hi - lo) * d[:, 0:3]
    dv = -1 + 2 * d[:, 3:6]
    ln = np.array([scenes.js_hypot(*r) or 1.0 for r in dv])
    tmax = tlo + (thi - tlo) * d[:, 6]
    return np.concatenate([o, dv / ln[:, None]], axis=1), tmax


@pytest.mark.parametrize("walk", [0, 1], ids=["ordered", "stackless"])
def test_walk_respects_finite_tmax(O, ps, walk):  # bvh.test.ts:114-131
    rays, tmax = _rays_finite(O, 5, 300, -12, 12, 0.5, 20)
    a = O.closest_batch(ps, rays, tmax, walk=walk)
    b = O.closest_batch(ps, rays, tmax, walk=O.WALK_BRUTE)
    assert (a["prim"] == b["prim"]).all()
    h = a["prim"] >= 0
    assert (a["t"][h] < tmax[h]).all()


def test_occluded_matches_brute(O, ps):  # bvh.test.ts:134-146
    rays, tmax = _rays_finite(O, 7, 500, -12, 12, 1, 30)
    occ, _ = O.occluded_batch(ps, rays, tmax)
    b = O.closest_batch(ps, rays, tmax, walk=O.WALK_BRUTE)
    assert (occ == (b["prim"] >= 0)).all()


def test_degenerate_inputs(O):  # bvh.test.ts:148-169
    empty = O.pack_scene({"objects": [], "sky": {"kind": "none"}})
    assert empty.primCount == 0 and empty.nodeCount == 1
    r = O.closest_batch(empty, np.array([[0, 0, 0, 0, 0, 1.0]]))
    assert r["prim"][0] == -1
    co = O.pack_scene(scenes.coincident_spheres(40))
    rep = O.validate_bvh(co.nodes, co.nodeCount, co.order, co.primCount)
    assert rep["errors"] == 0 and rep["maxLeafSize"] <= 4
    h = O.closest_batch(co, np.array([[1, 2, -5, 0, 0, 1.0]]))
    assert h["prim"][0] >= 0 and abs(h["t"][0] - 7.5) < 1e-5


def partition_closed_form(good):
    """Destination of each element of the in-place two-pointer partition
    (bvh.ts:262-278) from prefix counts alone; this is the rule the GPU builder
    applies.  Returns perm with perm[dest] = source index."""
    good = np.asarray(good, bool)
    n = len(good)
    bad_pos = np.flatnonzero(~good)                 # a_k: k-th bad from the left
    good_r = np.flatnonzero(good[::-1])             # b_k: k-th good from the right (R-index)
    kmax = min(len(bad_pos), len(good_r))
    K = 0
    while K < kmax and bad_pos[K] + good_r[K] + 2 <= n:
        K += 1
    bK = good_r[K - 1] if K > 0 else -1
    if K < len(bad_pos) and bad_pos[K] + bK + 2 <= n:
        x = bad_pos[K] + 1
    else:
        x = n - (bK + 1)
    perm = np.full(n, -1)
    badrank = np.cumsum(~good) - (~good)            # bads strictly before position
    goodrank_r = (np.cumsum(good[::-1]) - good[::-1])[::-1]  # goods strictly after position
    for i in range(n):
        if i < x:
            if good[i]:
                dest = i
            else:
                k = badrank[i]                      # 0-based rank among L-bads
                dest = n - 2 - (good_r[k - 1] if k > 0 else -1)
        else:
            r = n - 1 - i
            if good[i]:
                k = goodrank_r[i]
                dest = bad_pos[k]
            else:
                dest = n - 2 - r
        assert perm[dest] == -1
        perm[dest] = i
    return perm, int(good.sum())


def test_partition_closed_form_matches_two_pointer(O):
    rs = np.random.default_rng(7)
    for trial in range(4000):
        n = int(rs.integers(1, 41))
        p = rs.uniform(0.05, 0.95)
        good = rs.random(n) < p
        ref, lo = O.partition_flags(good.astype(np.uint8))
        mine, ng = partition_closed_form(good)
        assert lo == ng
        assert (ref == mine).all(), (good.astype(int), ref, mine)
