"""GPU parity: HIP kernels through the C ABI against the CPU oracle.

Integer/index results (closest-hit prim ids, BVH node arrays, order, occlusion
flags, traversal counters of the stackless walk) must be bit-exact.  Closest-hit
t/u/v are binary64 on both sides and compared exactly as well.  Images must
match per pixel within 1e-4 relative (the tolerance of the task statement);
the observed difference is printed.
"""
import math

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def rel_err(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    den = np.maximum(np.abs(g), np.abs(o))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(den > 0, np.abs(g - o) / den, 0.0)
    return r


@pytest.fixture(scope="module")
def bvh_scene(O):
    return O.pack_scene(scenes.random_scene(O.Rng(1234), 1000, 2000))


@pytest.fixture(scope="module")
def many_rays(O):
    return scenes.ray_set_bvh_test(O.Rng(99), 20000)


@pytest.mark.parametrize("walk", [0, 1, 2], ids=["stackless", "ordered", "wide"])
def test_closest_hit_bit_exact(ct, O, bvh_scene, many_rays, walk):
    owalk = O.WALK_STACKLESS if walk == ct.WALK_STACKLESS else O.WALK_ORDERED
    g = ct.closest_hit(bvh_scene, many_rays, walk=walk)
    o = O.closest_batch(bvh_scene, many_rays, walk=owalk)
    assert (g["prim"] == o["prim"]).all()
    assert np.array_equal(g["t"], o["t"])
    hit = o["prim"] >= 0
    if walk == ct.WALK_WIDE:
        # a sphere hit keeps u, v from the last triangle test (intersect.ts:173-176),
        # which depends on the visiting order: compare them on triangle hits
        kind = np.asarray(bvh_scene.prims, np.float32).reshape(-1, 16)[:, 3]
        hit &= kind[np.maximum(o["prim"], 0)] != 0
    assert np.array_equal(g["u"][hit], o["u"][hit]) and np.array_equal(g["v"][hit], o["v"][hit])
    assert hit.sum() > 5000


def test_closest_hit_finite_tmax_and_counts(ct, O, bvh_scene):
    rng = O.Rng(5)
    n = 5000
    d = rng.take(7 * n).reshape(n, 7)
    o = -12 + 24 * d[:, :3]
    dv = -1 + 2 * d[:, 3:6]
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    rays = np.concatenate([o, dv], 1)
    tmax = 0.5 + 19.5 * d[:, 6]
    for walk, owalk in ((ct.WALK_STACKLESS, O.WALK_STACKLESS), (ct.WALK_ORDERED, O.WALK_ORDERED),
                        (ct.WALK_WIDE, O.WALK_ORDERED)):
        g = ct.closest_hit(bvh_scene, rays, tmax, walk=walk)
        r = O.closest_batch(bvh_scene, rays, tmax, walk=owalk)
        assert (g["prim"] == r["prim"]).all() and np.array_equal(g["t"], r["t"])
    # traversal work: the stackless walk is an exact mirror, so are its counters
    gc = ct.closest_hit_counts(bvh_scene, rays, tmax, walk=ct.WALK_STACKLESS)
    oc = O.closest_batch(bvh_scene, rays, tmax, walk=O.WALK_STACKLESS)["counts"]
    assert gc == (oc[0], oc[1], oc[2])
    # the GPU's ordered walk prunes on float32-rounded stack entries and tests
    # parked leaves late: it visits a superset of the reference walk's nodes
    gc = ct.closest_hit_counts(bvh_scene, rays, tmax, walk=ct.WALK_ORDERED)
    oc = O.closest_batch(bvh_scene, rays, tmax, walk=O.WALK_ORDERED)["counts"]
    assert gc[0] >= oc[0] and gc[0] <= oc[0] * 1.5


def test_occluded_bit_exact(ct, O, bvh_scene):
    rng = O.Rng(7)
    n = 5000
    d = rng.take(7 * n).reshape(n, 7)
    o = -12 + 24 * d[:, :3]
    dv = -1 + 2 * d[:, 3:6]
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    rays = np.concatenate([o, dv], 1)
    tmax = 1 + 29 * d[:, 6]
    g = ct.occluded(bvh_scene, rays, tmax)
    r, _ = O.occluded_batch(bvh_scene, rays, tmax)
    assert (g == r).all() and 0 < g.sum() < n


def _translated(scene, off):
    """The scene moved by `off`: coordinates of large magnitude around small features."""
    out = dict(scene, objects=[])
    for ob in scene["objects"]:
        sh = dict(ob["shape"])
        for k in ("a", "b", "c", "center"):
            if k in sh:
                sh[k] = tuple(x + d for x, d in zip(sh[k], off))
        out["objects"].append(dict(ob, shape=sh))
    return out


def _aimed_rays(ps, seed, n):
    """Rays aimed at points of the scene's triangles -- vertices, edge midpoints
    and interior points, so many hits fall on shared boundaries -- from
    distances log-uniform in [1e-2, 1e5], a third of them with one or two
    direction components exactly zero."""
    rng = np.random.default_rng(seed)
    pr = np.asarray(ps.prims, np.float32).reshape(-1, 16).astype(np.float64)
    tri = np.nonzero(pr[:, 3] != 0)[0]
    k = tri[rng.integers(0, len(tri), n)]
    a, b, c = pr[k, 0:3], pr[k, 4:7], pr[k, 8:11]
    uv = rng.random((n, 2))
    uv = np.where(uv.sum(1, keepdims=True) > 1, 1 - uv, uv)
    fixed = np.array([[0, 0], [1, 0], [0, 1], [0.5, 0], [0, 0.5], [0.5, 0.5]])
    pick = rng.integers(0, len(fixed) + 4, n)
    uv = np.where((pick < len(fixed))[:, None], fixed[np.minimum(pick, len(fixed) - 1)], uv)
    target = a + uv[:, :1] * (b - a) + uv[:, 1:] * (c - a)
    d = rng.normal(size=(n, 3))
    zero = rng.integers(0, 9, n)
    d[zero == 0, 0] = 0
    d[zero == 1, 1] = 0
    d[zero == 2, 2] = 0
    d[zero == 3, :2] = 0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dist = 10.0 ** rng.uniform(-2, 5, n)
    return np.concatenate([target - dist[:, None] * d, d], 1)


@pytest.mark.parametrize("off", [(0, 0, 0), (4096.3, -3000.7, 65536.1)], ids=["origin", "far"])
def test_wide_walk_adversarial_rays(ct, O, off):
    """The 4-wide walk decides boxes in FP32 from the rounded binary64 ray; its
    per-axis slack must keep every box a binary64 hit lies in.  Rays from up
    to 1e5 away, axis-parallel rays and hits on vertices and edges, on the
    scene at the origin and moved to coordinates around 6.5e4, give the
    reference's closest hit (prim and t) exactly."""
    ps = O.pack_scene(_translated(scenes.random_scene(O.Rng(77), 50, 3000), off))
    rays = _aimed_rays(ps, 11, 20000)
    g = ct.closest_hit(ps, rays, walk=ct.WALK_WIDE)
    o = O.closest_batch(ps, rays, walk=O.WALK_ORDERED)
    assert (g["prim"] == o["prim"]).all()
    assert np.array_equal(g["t"], o["t"])
    assert (o["prim"] >= 0).mean() > 0.7


def test_scene_from_device_arrays(ct, O):
    """ct_scene_create_device: the packed arrays handed over in device memory
    give the scene the host arrays give (its 4-wide tree, primitive boxes
    included, is derived from a host copy): closest hits equal to the host
    scene's and the oracle's, the same image."""
    import ctypes as C
    A = ct._capi
    sc = scenes.random_scene(O.Rng(77), 50, 3000)
    sc["objects"].append(scenes.obj("light", scenes.sphere((0, 12, 0), 2.0),
                                    scenes.light(scenes.rgb(1, 1, 1), 10)))
    ps = O.pack_scene(sc)
    host = ct.DeviceScene(ps)
    arrays = [ct.DeviceBuffer.from_array(np.ascontiguousarray(a, t)) for a, t in
              ((ps.prims, np.float32), (ps.nodes, np.float32), (ps.mats, np.float32),
               (ps.lights, np.int32), (ps.primObject, np.uint32))]
    sky = A.Sky()
    h = C.c_void_p()
    A.check(A.lib().ct_scene_create_device(
        arrays[0].ptr, ps.primCount, arrays[1].ptr, ps.nodeCount, arrays[2].ptr, ps.matCount,
        arrays[3].ptr, len(ps.lights), arrays[4].ptr, C.byref(sky), None, C.byref(h)),
        "ct_scene_create_device")
    dev = object.__new__(ct.DeviceScene)
    dev.handle, dev.primCount, dev.nodeCount = h, ps.primCount, ps.nodeCount
    rays = _aimed_rays(ps, 13, 4000)
    gd = ct.closest_hit(dev, rays, walk=ct.WALK_WIDE)
    gh = ct.closest_hit(host, rays, walk=ct.WALK_WIDE)
    o = O.closest_batch(ps, rays, walk=O.WALK_ORDERED)
    assert np.array_equal(gd["prim"], gh["prim"]) and np.array_equal(gd["prim"], o["prim"])
    assert np.array_equal(gd["t"], o["t"])
    fr = O.camera_frame(sc["camera"], 1.5)
    opts = {"width": 48, "height": 32, "spp": 2, "seed": 4, "maxBounces": 6}
    imgs = []
    for s_ in (dev, host):
        d = ct.DeviceBuffer.empty((48 * 32 * 3,), np.float32)
        tr = ct.Tracer()
        tr.set_path_kernel(ct.PATHS_WAVEFRONT)
        tr.render_image(s_, fr, opts, d)
        imgs.append(d.numpy())
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    assert (imgs[0] > 0).mean() > 0.1
    dev.close()


def test_leaves_of_up_to_seven_primitives(ct, O):
    """Leaf words hold up to seven primitives (bvh.ts:60-68); the reference
    builder stops at four, a packed tree from elsewhere need not.  A tree
    whose leaf holds six scattered primitives: the 4-wide tree derived at
    upload, which may give that leaf's primitives up to four boxes of runs,
    returns the reference walks' closest hits."""
    rng = np.random.default_rng(12)
    grey = scenes.lambert(scenes.rgb(0.5, 0.5, 0.5))
    objs = []
    for i in range(7):
        c = rng.uniform(-2, 2, 3)
        a, b, d = (tuple(c + rng.uniform(-0.15, 0.15, 3)) for _ in range(3))
        objs.append(scenes.obj(f"t{i}", {"kind": "triangle", "a": a, "b": b, "c": d}, grey))
    base = O.pack_scene({"objects": objs, "sky": {"kind": "none"}})
    prims = np.asarray(base.prims, np.float32).copy()
    pr = prims.reshape(-1, 16)[:7].astype(np.float64)
    nodes = np.zeros(3 * 8, np.float32)
    for n, ix, skip, leaf in ((0, range(7), 3, -1), (1, range(6), 2, 0 * 8 + 6),
                              (2, [6], 3, 6 * 8 + 1)):
        v = np.concatenate([pr[list(ix), k:k + 3] for k in (0, 4, 8)])
        nodes[n * 8:n * 8 + 3] = v.min(0) - 1e-3
        nodes[n * 8 + 3] = skip
        nodes[n * 8 + 4:n * 8 + 7] = v.max(0) + 1e-3
        nodes[n * 8 + 7] = leaf
    ps = O.Packed(prims, nodes, 3, np.asarray(base.mats, np.float32).copy(),
                  np.asarray(base.lights, np.int32).copy(),
                  np.asarray(base.primObject, np.uint32).copy(), np.asarray(base.order).copy(),
                  {"kind": "none"})
    rays = _aimed_rays(ps, 21, 3000)
    for walk, owalk in ((ct.WALK_WIDE, O.WALK_ORDERED), (ct.WALK_ORDERED, O.WALK_ORDERED),
                        (ct.WALK_STACKLESS, O.WALK_STACKLESS)):
        g = ct.closest_hit(ps, rays, walk=walk)
        o = O.closest_batch(ps, rays, walk=owalk)
        assert np.array_equal(g["prim"], o["prim"]) and np.array_equal(g["t"], o["t"]), walk
    assert (o["prim"] >= 0).mean() > 0.3


@pytest.mark.parametrize("preset", ["cornell", "mirror"])
def test_exact_ties_follow_the_reference_walks(ct, O, preset):
    """On exact-t ties the ordered walk (intersect.ts:202-282) keeps the
    primitive of the leaf it reaches first (nearer child first, tl <= tr to the
    left) and the stackless walk (intersect.ts:290-365) the lowest packed
    index.  The GPU's ordered and 4-wide walks must return the former, its
    stackless walk the latter."""
    sc = ct.presetById(preset)["scene"]
    ps = O.pack_scene(sc)
    rays = scenes.tie_rays(ps, 3, 12 if preset == "cornell" else 2)
    ordered = O.closest_batch(ps, rays, walk=O.WALK_ORDERED)
    stackless = O.closest_batch(ps, rays, walk=O.WALK_STACKLESS)
    ties = int((ordered["prim"] != stackless["prim"]).sum())
    assert ties > 0  # the ray set does produce ties
    for walk, ref in ((ct.WALK_ORDERED, ordered), (ct.WALK_WIDE, ordered),
                      (ct.WALK_STACKLESS, stackless)):
        g = ct.closest_hit(ps, rays, walk=walk)
        bad = np.flatnonzero(g["prim"] != ref["prim"])
        assert len(bad) == 0, (walk, ties, bad[:8], g["prim"][bad[:8]], ref["prim"][bad[:8]])
        assert np.array_equal(g["t"], ref["t"])


def test_empty_and_coincident(ct, O):
    empty = O.pack_scene({"objects": [], "sky": {"kind": "none"}})
    g = ct.closest_hit(empty, np.array([[0, 0, 0, 0, 0, 1.0]]))
    assert g["prim"][0] == -1 and g["t"][0] == math.inf
    co = O.pack_scene(scenes.coincident_spheres(40))
    g = ct.closest_hit(co, np.array([[1, 2, -5, 0, 0, 1.0]]))
    r = O.closest_batch(co, np.array([[1, 2, -5, 0, 0, 1.0]]))
    assert g["prim"][0] == r["prim"][0] and g["t"][0] == r["t"][0]
    assert abs(g["t"][0] - 7.5) < 1e-5


# ---------------------------------------------------------------------------
# BVH build

def _bounds_inputs(O, scene):
    kind, geom, obj = O.expand_scene(scene)
    return O.prim_bounds(kind, geom)


BUILD_CASES = {
    "bvh_test_scene": lambda O: scenes.random_scene(O.Rng(1234), 1000, 2000),
    "cornell": lambda O: scenes.cornell_box(),
    "coincident40": lambda O: scenes.coincident_spheres(40),
    "single": lambda O: scenes.coincident_spheres(1),
    "five": lambda O: scenes.random_scene(O.Rng(3), 2, 3),
    "glass_spheres": lambda O: scenes.glass_spheres(O.Rng(20240607)),
}


@pytest.mark.parametrize("case", sorted(BUILD_CASES))
@pytest.mark.parametrize("bins", [12, 16])
def test_gpu_bvh_build_bit_exact(ct, O, case, bins):
    sc = BUILD_CASES[case](O)
    bmin, bmax, bcen = _bounds_inputs(O, sc)
    on, onc, oo = O.build_bvh(bmin, bmax, bcen, bins)
    gn, gnc, go = ct.buildBvh(bmin.reshape(-1, 3), bmax.reshape(-1, 3), bcen.reshape(-1, 3), bins)
    assert gnc == onc
    assert np.array_equal(gn.view(np.uint32), on.view(np.uint32))
    assert np.array_equal(go, oo)


def test_gpu_bvh_build_hand_derived_fixtures(ct):
    """The GPU builder against node arrays derived by hand from bvh.ts's rules
    (tests/golden/make_bvh_fixtures.py): equal-cost splits, axis ties, the
    partition's reordering and the median fallback."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "bvh_fixtures.json")) as f:
        fixtures = json.load(f)
    for fx in fixtures:
        b = [np.array(fx[k]).reshape(-1, 3) for k in ("bmin", "bmax", "bcen")]
        gn, gnc, go = ct.buildBvh(*b, fx["bins"])
        assert gnc == fx["nodeCount"], fx["name"]
        assert np.array_equal(gn.view(np.uint32), np.array(fx["nodes_u32"], np.uint32)), fx["name"]
        assert list(go) == fx["order"], fx["name"]


def test_gpu_bvh_build_empty(ct):
    z = np.zeros((0, 3))
    nodes, nc, order = ct.buildBvh(z, z, z)
    assert nc == 1 and list(nodes) == [0, 0, 0, 1, 0, 0, 0, 0] and len(order) == 0


@pytest.mark.parametrize("n", [20000, 200000])
def test_gpu_bvh_build_soup_bit_exact(ct, O, n):
    syn = ct.synth
    rng = np.random.default_rng(n)
    tris = syn.equilateral_soup(rng, n)
    kind = np.ones(n, np.int32)
    geom = tris.reshape(n, 9)
    bmin, bmax, bcen = O.prim_bounds(kind, geom)
    on, onc, oo = O.build_bvh(bmin, bmax, bcen, 12)
    gn, gnc, go = ct.buildBvh(bmin.reshape(-1, 3), bmax.reshape(-1, 3), bcen.reshape(-1, 3), 12)
    assert gnc == onc
    assert np.array_equal(gn.view(np.uint32), on.view(np.uint32))
    assert np.array_equal(go, oo)


def _degenerate_bounds(case):
    """PrimBounds whose nodes fall back to medianSplit (bvh.ts:284-296) with more
    primitives than one workgroup sorts in LDS."""
    rs = np.random.default_rng(31)
    if case == "coincident5000":  # one centroid: the stable sort keeps position order
        n = 5000
        cen = np.tile([1.0, 2.0, 3.0], (n, 1))
    elif case == "ties_x6000":  # centroid x in 1e-16 steps (extent <= 1e-12), many ties
        n = 6000
        cen = np.zeros((n, 3))
        cen[:, 0] = 0.25 + rs.integers(0, 9000, n) * 1e-16
        cen[:, 1] = -1.0 + rs.integers(0, 3, n) * 1e-16
        cen[:, 2] = 4.0
    elif case == "ties_y4100":  # the widest degenerate axis is y
        n = 4100
        cen = np.zeros((n, 3))
        cen[:, 0] = 3.0 + rs.integers(0, 2, n) * 1e-16
        cen[:, 1] = rs.integers(-4000, 4000, n) * 1e-16
        cen[:, 2] = -7.0 + rs.integers(0, 5, n) * 1e-17
    else:  # "mixed": a SAH top over a 3000-primitive cluster and a spread cloud
        n = 6000
        cen = np.zeros((n, 3))
        cen[:3000] = [0.5, 0.5, 0.5]
        cen[3000:] = rs.uniform(-10, 10, (3000, 3))
        cen = cen[rs.permutation(n)]
    half = rs.uniform(0.01, 1.0, (n, 3))
    return cen - half, cen + half, cen


@pytest.mark.parametrize("case", ["coincident5000", "ties_x6000", "ties_y4100", "mixed"])
def test_gpu_bvh_build_large_median_nodes(ct, O, case):
    bmin, bmax, bcen = _degenerate_bounds(case)
    on, onc, oo = O.build_bvh(bmin.ravel(), bmax.ravel(), bcen.ravel(), 12)
    rep = O.validate_bvh(on, onc, oo, len(bcen))
    assert rep["errors"] == 0
    gn, gnc, go = ct.buildBvh(bmin, bmax, bcen, 12)
    assert gnc == onc
    assert np.array_equal(gn.view(np.uint32), on.view(np.uint32))
    assert np.array_equal(go, oo)


def test_pack_scene_matches_oracle(ct, O):
    sc = scenes.random_scene(O.Rng(1234), 1000, 2000)
    g = ct.packScene(sc)
    o = O.pack_scene(sc)
    assert g.primCount == o.primCount and g.nodeCount == o.nodeCount
    assert np.array_equal(g.prims.view(np.uint32), o.prims.view(np.uint32))
    assert np.array_equal(g.nodes.view(np.uint32), o.nodes.view(np.uint32))
    assert np.array_equal(g.mats.view(np.uint32), o.mats.view(np.uint32))
    assert np.array_equal(g.lights, o.lights) and np.array_equal(g.primObject, o.primObject)
    for mk in (scenes.cornell_box, scenes.night_emissives):
        sc = mk()
        g, o = ct.packScene(sc), O.pack_scene(sc)
        assert np.array_equal(g.prims.view(np.uint32), o.prims.view(np.uint32))
        assert np.array_equal(g.nodes.view(np.uint32), o.nodes.view(np.uint32))
        assert np.array_equal(g.lights, o.lights)


# ---------------------------------------------------------------------------
# rendering

def _render_pair(ct, O, sc, w, h, spp, seed, **kw):
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], w / h)
    o = O.render_image(ps, fr, w, h, spp, seed=seed, **kw)
    opts = {"width": w, "height": h, "spp": spp, "seed": seed,
            "maxBounces": kw.get("max_bounces", 8),
            "russianRoulette": kw.get("russian_roulette", True), "nee": kw.get("nee", True)}
    g = ct.renderImage(ps, fr, opts)
    return g, o


def test_render_cornell_matches_oracle_and_golden(ct, O):
    from test_oracle_golden import GOLDEN, read_ppm
    g, o = _render_pair(ct, O, scenes.cornell_box(), 96, 64, 16, 1)
    err = rel_err(g, o)
    print(f"cornell 96x64x16: max rel err {err.max():.3e}, exact pixels {(g == o).mean():.4f}")
    assert err.max() <= REL_TOL
    rgba = O.tonemap_buffer(g, 96, 64, 1, 0.0, "aces").reshape(64, 96, 4)
    _, _, gold = read_ppm(GOLDEN)
    diff = np.abs(rgba[:, :, :3].astype(int) - gold.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3


@pytest.mark.parametrize("walk", [0, 1, 2], ids=["stackless", "ordered", "wide"])
def test_render_c1_cornell_variant(ct, O, walk):
    sc = scenes.cornell_c1()
    ps = O.pack_scene(sc, bins=16)
    w = h = 128
    fr = O.camera_frame(sc["camera"], 1.0)
    o = O.render_image(ps, fr, w, h, 4, seed=1, max_bounces=8)
    tr = ct.Tracer(walk=walk)
    tr.set_traversal(walk, 16)
    tr.set_path_kernel(ct.PATHS_WAVEFRONT)  # the walks of the wavefront launches
    d = ct.DeviceBuffer.empty((w * h * 3,), np.float32)
    tr.render_image(ps, fr, {"width": w, "height": h, "spp": 4, "seed": 1, "maxBounces": 8}, d)
    g = d.numpy()
    err = rel_err(g, o)
    print(f"C1 128^2x4 walk={walk}: max rel err {err.max():.3e}")
    assert err.max() <= REL_TOL


def test_accumulate_pass_lens_sky_checker(ct, O):
    sc = scenes.glass_spheres(O.Rng(20240607))
    ps = O.pack_scene(sc)
    w, h = 96, 64
    fr = O.camera_frame(sc["camera"], w / h)
    acc_o = np.zeros(w * h * 3, np.float32)
    acc_g = np.zeros(w * h * 3, np.float32)
    for f in (11, 12):
        O.accumulate_pass(ps, fr, w, h, f, acc_o, max_bounces=8)
        ct.accumulatePass(ps, fr, w, h, f, {"maxBounces": 8}, acc_g)
    err = rel_err(acc_g, acc_o)
    print(f"glass spheres 2 passes: max rel err {err.max():.3e}")
    assert err.max() <= REL_TOL


def test_sphere_lights(ct, O):
    g, o = _render_pair(ct, O, scenes.night_emissives(), 64, 48, 4, 2)
    err = rel_err(g, o)
    print(f"night emissives: max rel err {err.max():.3e}")
    assert err.max() <= REL_TOL


@pytest.mark.parametrize("nee,rr,seed,tol_mean,tol_px", [(False, False, 3, 0.05, 0.05),
                                                         (True, False, 9, 0.03, 0.08),
                                                         (False, True, 5, 0.03, None)])
def test_furnace_on_gpu(ct, O, nee, rr, seed, tol_mean, tol_px):
    sc = scenes.furnace()
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 1.0)
    w = 8 if not nee and not rr else 4
    img = ct.renderImage(ps, fr, {"width": w, "height": w, "spp": 4096, "seed": seed,
                                  "maxBounces": 64, "russianRoulette": rr, "nee": nee})
    assert abs(img.mean() - 5) / 5 < tol_mean
    if tol_px is not None:
        assert np.all(np.abs(img - 5) / 5 < tol_px)


def test_render_is_deterministic(ct, O):
    sc = scenes.cornell_c1()
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 1.0)
    opts = {"width": 64, "height": 64, "spp": 8, "seed": 4, "maxBounces": 8}
    a = ct.renderImage(ps, fr, opts)
    b = ct.renderImage(ps, fr, opts)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("walk,stack", [(2, 16), (2, 8), (1, 8)], ids=["wide16", "wide8", "ordered8"])
def test_render_soup_all_walks(ct, O, walk, stack):
    """20k-triangle soup (materials of all kinds, sky, 64 lights): the wide and
    binary ordered walks, with stacks small enough to force the stackless
    fallback, against the oracle."""
    sc = ct.synth.soup_scene(20000, n_lights=64, n_materials=50, seed=5)
    ps = O.pack_scene(sc)
    w, h = 96, 64
    fr = O.camera_frame(sc["camera"], w / h)
    o = O.render_image(ps, fr, w, h, 2, seed=3, max_bounces=6)
    tr = ct.Tracer(walk=walk)
    tr.set_traversal(walk, stack)
    tr.set_path_kernel(ct.PATHS_WAVEFRONT)  # the walks of the wavefront launches
    tr.set_counting(True)
    d = ct.DeviceBuffer.empty((w * h * 3,), np.float32)
    tr.render_image(ps, fr, {"width": w, "height": h, "spp": 2, "seed": 3, "maxBounces": 6}, d)
    g = d.numpy()
    st = tr.stats()
    err = rel_err(g, o)
    print(f"soup 20k walk={walk} stack={stack}: max rel err {err.max():.3e}, "
          f"fallbacks {st['stackFallbacks']}/{st['shadowStackFallbacks']}")
    assert err.max() <= REL_TOL


@pytest.mark.parametrize("preset", ["cornell", "glass", "night", "dof", "mirror"])
def test_presets_gpu_bvh_and_render(ct, O, preset):
    """The five presets of scenes.ts:45-237 (lens draws in Depth of Field, the
    torus-knot mesh and metal walls of Mirror Room, sphere lights without a sky
    in Night Emissives): the GPU builder reproduces the reference BVH, and a
    render through it matches the oracle per pixel within 1e-4."""
    entry = next(p for p in ct.presets.SCENE_PRESETS if p["id"] == preset)
    sc = entry["scene"]
    o_ps = O.pack_scene(sc)
    g_ps = ct.packScene(sc)
    assert g_ps.nodeCount == o_ps.nodeCount
    assert np.array_equal(g_ps.nodes.view(np.uint32), o_ps.nodes.view(np.uint32))
    assert np.array_equal(g_ps.prims.view(np.uint32), o_ps.prims.view(np.uint32))
    w, h, spp = 48, 32, 2
    fr = O.camera_frame(sc["camera"], w / h)
    o = O.render_image(o_ps, fr, w, h, spp, seed=3, max_bounces=8)
    g = ct.renderImage(g_ps, fr, {"width": w, "height": h, "spp": spp, "seed": 3,
                                  "maxBounces": 8})
    err = rel_err(g, o)
    print(f"preset {preset} {w}x{h}x{spp}: max rel err {err.max():.3e}, "
          f"exact pixels {(g == o).mean():.4f}")
    assert np.isfinite(g).all() and g.max() > 0
    assert err.max() <= REL_TOL


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_mirror_paths_bit_identical(ct, O, seed):
    """Chains of rough-metal bounces (Mirror Room: cbrt-scaled fuzz, cosine and
    sphere directions from sin/cos of 2*pi*u) -- the GPU's Math.sin/cos/cbrt
    are the engine's, bit for bit, so every pixel equals the oracle's exactly,
    not only within 1e-4."""
    sc = next(p for p in ct.presets.SCENE_PRESETS if p["id"] == "mirror")["scene"]
    ps = O.pack_scene(sc)
    w, h, spp = 40, 30, 2
    fr = O.camera_frame(sc["camera"], w / h)
    g = ct.renderImage(ps, fr, {"width": w, "height": h, "spp": spp, "seed": seed,
                                "maxBounces": 6})
    o = O.render_image(ps, fr, w, h, spp, seed=seed, max_bounces=6)
    bad = np.flatnonzero(g.view(np.uint32) != o.view(np.uint32))
    assert len(bad) == 0, (bad[:6] // 3, g[bad[:6]], o[bad[:6]])


@pytest.mark.parametrize("preset", ["cornell", "glass", "night", "dof", "mirror"])
def test_ingested_scene_renders_like_the_oracle(ct, O, preset):
    """Scene ingestion end to end: the preset written as a caustic-scene
    document (sceneToJSON, serialize.ts:128-139), parsed back (parseSceneJSON,
    serialize.ts:172-179), packed with the GPU builder and rendered on the GPU
    -- arrays identical to the oracle's pack of the original preset, image
    within 1e-4 of the oracle's render of it."""
    sc = next(p for p in ct.presets.SCENE_PRESETS if p["id"] == preset)["scene"]
    doc = ct.sceneToJSON(sc)
    parsed = ct.parseSceneJSON(doc)
    g_ps = ct.packScene(parsed)
    o_ps = O.pack_scene(sc)
    for k in ("prims", "nodes", "mats"):
        assert np.array_equal(getattr(g_ps, k).view(np.uint32), getattr(o_ps, k).view(np.uint32)), k
    assert np.array_equal(g_ps.lights, o_ps.lights)
    assert np.array_equal(g_ps.primObject, o_ps.primObject)
    w, h, spp = 40, 30, 2
    fr = ct.cameraFrame(parsed["camera"], w / h)
    assert np.array_equal(fr, O.camera_frame(sc["camera"], w / h))
    g = ct.renderImage(g_ps, fr, {"width": w, "height": h, "spp": spp, "seed": 11,
                                  "maxBounces": 6})
    o = O.render_image(o_ps, fr, w, h, spp, seed=11, max_bounces=6)
    assert np.isfinite(g).all() and g.max() > 0
    assert rel_err(g, o).max() <= REL_TOL


@pytest.mark.parametrize("w,h,spp,depth", [(1, 1, 1, 0), (3, 1, 2, 1), (1, 5, 3, 0), (7, 3, 1, 2),
                                           (2, 2, 0, 3)])
def test_render_small_shapes_and_depths(ct, O, w, h, spp, depth):
    """Single-pixel, single-row and single-column images, zero and one bounce,
    and zero samples (the reference leaves NaN: zero sums times 1/0): the GPU
    image equals the oracle's, NaN for NaN, within the per-pixel tolerance."""
    sc = scenes.cornell_box()
    ps = O.pack_scene(sc)
    frame = O.camera_frame(sc["camera"], w / h)
    g = ct.renderImage(ps, frame, {"width": w, "height": h, "spp": spp, "seed": 5,
                                   "maxBounces": depth})
    o = O.render_image(ps, frame, w, h, spp, seed=5, max_bounces=depth)
    assert g.shape == o.shape == (w * h * 3,)
    assert np.array_equal(np.isnan(g), np.isnan(o))
    if spp == 0:
        assert np.isnan(g).all()
    else:
        assert np.isfinite(g).all()
        assert rel_err(g, o).max() <= REL_TOL


def test_deepest_paths_and_the_depth_limit(ct, O):
    """maxBounces 254, the deepest the per-bounce counters hold: without Russian
    roulette every path in the closed furnace bounces 254 times, and the image
    equals the oracle's; 255 is refused with CT_ELIMIT (tracer.ts:209 loops to
    any depth), never traced to a different depth."""
    sc = scenes.furnace()
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 4 / 3)
    opts = {"width": 4, "height": 3, "spp": 2, "seed": 3, "maxBounces": 254,
            "russianRoulette": False, "nee": False}
    g = ct.renderImage(ps, fr, opts)
    o = O.render_image(ps, fr, 4, 3, 2, seed=3, max_bounces=254, russian_roulette=False,
                       nee=False)
    # every path sees the shell's emission at each of its 255 hits: sum of 0.8^k
    assert np.allclose(g, 5 * (1 - 0.8 ** 255), rtol=1e-6)
    assert rel_err(g, o).max() <= REL_TOL
    with pytest.raises(Exception, match="maxBounces"):
        ct.renderImage(ps, fr, dict(opts, maxBounces=255))


@pytest.mark.parametrize("sky", [{"kind": "none"},
                                 {"kind": "gradient", "zenith": (0.3, 0.5, 0.9),
                                  "horizon": (1.0, 0.9, 0.8), "strength": 1.5}],
                         ids=["black", "gradient"])
def test_render_empty_scene(ct, O, sky):
    """No primitives (bvh.ts:93-100's single empty leaf): every camera ray
    misses, the image is the sky of tracer.ts:211-216 or black, as the oracle's."""
    sc = {"objects": [], "sky": sky, "camera": scenes.cornell_box()["camera"]}
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 1.5)
    g = ct.renderImage(ps, fr, {"width": 12, "height": 8, "spp": 3, "seed": 2, "maxBounces": 4})
    o = O.render_image(ps, fr, 12, 8, 3, seed=2, max_bounces=4)
    assert np.array_equal(g, o)
    assert (g > 0).all() if sky["kind"] == "gradient" else (g == 0).all()
    assert not ct.occluded(ps, np.array([[0, 0, 0, 0, 1.0, 0]]), np.array([10.0]))[0]


@pytest.mark.parametrize("preset", ["cornell", "night", "glass", "mirror"])
def test_radiance_into_matches_oracle(ct, O, preset):
    """radianceInto (tracer.ts:191-328) of rays with the caller's RNG states:
    the radiance bit for bit and the RNG states after the calls equal the
    oracle's, for camera rays with pixel-seeded states and for rays from
    inside the scene with arbitrary states;
    also with NEE and Russian roulette off, and for rays past the path
    capacity (further jobs of one plan)."""
    sc = ct.presetById(preset)["scene"]
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 1.0)
    rs = np.random.default_rng(17)
    rays, states = [], []
    for k in range(200):
        x, y = int(rs.integers(0, 64)), int(rs.integers(0, 64))
        o, d = O.camera_sample(fr, 64, 64, x, y, k)
        rays.append(np.concatenate([o, d]))
        states.append(O.seed_for(y * 64 + x, 1000 + k))
    for k in range(100):
        o = rs.uniform(-0.5, 0.5, 3)
        v = rs.normal(size=3)
        rays.append(np.concatenate([o, v / np.linalg.norm(v)]))
        states.append(int(rs.integers(1, 1 << 32)))
    rays = np.array(rays)
    states = np.array(states, np.uint32)
    for opts in ({"maxBounces": 8}, {"maxBounces": 12, "nee": False, "russianRoulette": False}):
        L, st = ct.radiance_batch(ps, rays, states, opts)
        for i in range(len(rays)):
            oL, ost = O.radiance(ps, rays[i, :3], rays[i, 3:], int(states[i]),
                                 max_bounces=opts["maxBounces"],
                                 russian_roulette=opts.get("russianRoulette", True),
                                 nee=opts.get("nee", True), with_state=True)
            assert np.array_equal(L[i], oL), (i, L[i], oL)
            assert st[i] == ost
    # the single-ray mirror updates the caller's Rng as the reference does
    rng = ct.Rng.fromPixel(5, 9)
    out = ct.radianceInto(ps, *rays[0], rng, {"maxBounces": 8})
    oL, ost = O.radiance(ps, rays[0, :3], rays[0, 3:], int(ct.seedFor(5, 9)), with_state=True)
    assert np.array_equal(out, oL) and rng.state == ost
    # more rays than one round holds: a tracer of small capacity runs several jobs
    big = np.tile(rays, (9, 1))[:2500]
    big_states = (np.tile(states, 9)[:2500] + np.arange(2500, dtype=np.uint32)) | 1
    tr = ct.Tracer()
    tr.set_capacity(1024)  # three jobs of at most 1024 rays
    d_rays = ct.DeviceBuffer.from_array(np.ascontiguousarray(big))
    d_rng = ct.DeviceBuffer.from_array(big_states.copy())
    d_out = ct.DeviceBuffer.empty((len(big) * 3,), np.float64)
    tr.radiance(ps, d_rays, len(big), d_rng, {"maxBounces": 8}, d_out)
    L8, st8 = ct.radiance_batch(ps, big, big_states, {"maxBounces": 8})
    assert np.array_equal(d_out.numpy().reshape(-1, 3), L8)
    assert np.array_equal(d_rng.numpy(), st8)
