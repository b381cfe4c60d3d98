"""The per-path launch (k_paths: one lane carries a path through all its
bounces) against the wavefront launches and the oracle.

Both forms run the same binary64 arithmetic in the reference's order, so the
images, radiance values, RNG states and ray counts must be identical bit for
bit under every option, and within 1e-4 per pixel of the oracle (in practice
equal); CT_PATHS_AUTO takes the per-path launch for rounds of at most 2^24
paths over scenes of at most 32768 primitives, and the wavefront otherwise."""
import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def rel_err(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    den = np.maximum(np.abs(g), np.abs(o))
    return np.where(den > 0, np.abs(g - o) / np.where(den > 0, den, 1), 0.0)


def _render(ct, ps, fr, opts, mode, graph=False, counting=False, capacity=None, walk=None):
    tr = ct.Tracer() if walk is None else ct.Tracer(walk=walk)
    tr.set_path_kernel(mode)
    if graph:
        tr.set_graph(True)
    if counting:
        tr.set_counting(True)
    if capacity:
        tr.set_capacity(capacity)
    d = ct.DeviceBuffer.empty((opts["width"] * opts["height"] * 3,), np.float32)
    tr.render_image(ps, fr, opts, d)
    return d.numpy(), tr.stats(), tr.get_path_kernel()[1], tr


def _scene(ct, O, name):
    if name == "c1":
        sc = scenes.cornell_c1()
        return sc, O.pack_scene(sc, bins=16)
    if name == "glass":  # thin lens, gradient sky, checker ground, 60 spheres
        sc = scenes.glass_spheres(O.Rng(20240607))
    elif name == "night":  # sphere lights
        sc = scenes.night_emissives()
    elif name == "furnace":
        sc = scenes.furnace()
    elif name in ("mirror", "dof"):
        sc = ct.presetById(name)["scene"]
    else:
        sc = scenes.cornell_box()
    return sc, O.pack_scene(sc)


CASES = [
    ("c1", 64, 64, 4, 1, {"maxBounces": 8}),
    ("cornell", 48, 32, 8, 3, {"maxBounces": 8, "nee": False}),
    ("cornell", 40, 30, 4, 5, {"maxBounces": 12, "russianRoulette": False}),
    ("cornell", 32, 24, 2, 9, {"maxBounces": 0}),
    ("glass", 64, 48, 4, 2, {"maxBounces": 8}),
    ("night", 48, 36, 4, 6, {"maxBounces": 6}),
    ("dof", 48, 32, 3, 4, {"maxBounces": 8}),
    ("mirror", 40, 30, 2, 7, {"maxBounces": 254, "russianRoulette": False}),
    ("furnace", 8, 8, 64, 1, {"maxBounces": 64}),
]


@pytest.mark.parametrize("name,w,h,spp,seed,extra", CASES,
                         ids=[f"{c[0]}-{i}" for i, c in enumerate(CASES)])
def test_per_path_equals_wavefront_and_oracle(ct, O, name, w, h, spp, seed, extra):
    sc, ps = _scene(ct, O, name)
    fr = O.camera_frame(sc["camera"], w / h)
    opts = dict({"width": w, "height": h, "spp": spp, "seed": seed, "russianRoulette": True,
                 "nee": True}, **extra)
    gp, sp, rp, _ = _render(ct, ps, fr, opts, ct.PATHS_PER_PATH)
    gw, sw, rw, _ = _render(ct, ps, fr, opts, ct.PATHS_WAVEFRONT)
    assert rp == 1 and rw == 0
    assert np.array_equal(gp.view(np.uint32), gw.view(np.uint32))
    assert sp["extensionRays"] == sw["extensionRays"] and sp["shadowRays"] == sw["shadowRays"]
    assert sp["extensionRays"] > 0
    o = O.render_image(ps, fr, w, h, spp, seed=seed, max_bounces=opts["maxBounces"],
                       russian_roulette=opts["russianRoulette"], nee=opts["nee"])
    err = rel_err(gp, o)
    print(f"{name} {w}x{h}x{spp}: per-path = wavefront, max rel err vs oracle {err.max():.2e}, "
          f"exact {(gp == o).mean():.4f}, rays {sp['extensionRays']}+{sp['shadowRays']}")
    assert err.max() <= REL_TOL


def test_two_path_groups_and_several_rounds(ct, O):
    """A round of 2^20 paths (one per-path launch; the wavefront traces it as
    two path groups on two streams) and the same job in rounds of 2^18 paths:
    the per-path launches give the wavefront's image."""
    sc, ps = _scene(ct, O, "c1")
    fr = O.camera_frame(sc["camera"], 1.0)
    opts = {"width": 256, "height": 256, "spp": 16, "seed": 1, "maxBounces": 8}
    gw, sw, _, _ = _render(ct, ps, fr, opts, ct.PATHS_WAVEFRONT)
    gp, sp, rp, _ = _render(ct, ps, fr, opts, ct.PATHS_PER_PATH)
    gr, sr, rr, _ = _render(ct, ps, fr, opts, ct.PATHS_PER_PATH, capacity=1 << 18)
    assert rp == 1 and rr == 4
    for g, s in ((gp, sp), (gr, sr)):
        assert np.array_equal(g.view(np.uint32), gw.view(np.uint32))
        assert s["extensionRays"] == sw["extensionRays"] and s["shadowRays"] == sw["shadowRays"]


def test_auto_picks_per_path_for_small_scenes(ct, O):
    small = ct.synth.soup_scene(32700, n_lights=16, n_materials=8, seed=3)  # 32716 primitives
    large = ct.synth.soup_scene(32800, n_lights=16, n_materials=8, seed=3)  # 32816 primitives
    for sc, (w, h), want in ((small, (48, 32), 1), (large, (48, 32), 0),
                             (small, (1024, 1024), 1)):  # 3 x 2^20 paths, two path groups
        ps = ct.packScene(sc)
        fr = ct.cameraFrame(sc["camera"], w / h)
        opts = {"width": w, "height": h, "spp": 2 if w < 1024 else 3, "seed": 5, "maxBounces": 6}
        ga, _, ra, tr = _render(ct, ps, fr, opts, ct.PATHS_AUTO)
        assert tr.get_path_kernel() == (ct.PATHS_AUTO, want)
        gw, _, _, _ = _render(ct, ps, fr, opts, ct.PATHS_WAVEFRONT)
        assert np.array_equal(ga.view(np.uint32), gw.view(np.uint32))
    # rounds above 2^24 paths run the wavefront, whatever the scene
    sc, ps = _scene(ct, O, "cornell")
    opts = {"width": 1024, "height": 1024, "spp": 17, "seed": 3, "maxBounces": 4}
    fr = O.camera_frame(sc["camera"], 1.0)
    ga, _, ra, _ = _render(ct, ps, fr, opts, ct.PATHS_AUTO)
    gp, _, rp, _ = _render(ct, ps, fr, opts, ct.PATHS_PER_PATH)
    assert ra == 0 and rp == 1
    assert np.array_equal(ga.view(np.uint32), gp.view(np.uint32))
    # the FP32 precision form always runs its wavefront
    tr = ct.Tracer()
    tr.set_path_kernel(ct.PATHS_PER_PATH)
    tr.set_precision(ct.PRECISION_F32)
    d = ct.DeviceBuffer.empty((48 * 32 * 3,), np.float32)
    tr.render_image(ct.packScene(small), ct.cameraFrame(small["camera"], 1.5),
                    {"width": 48, "height": 32, "spp": 1, "seed": 5, "maxBounces": 6}, d)
    assert tr.get_path_kernel() == (ct.PATHS_PER_PATH, 0)
    with pytest.raises(ct.CTError):
        tr.set_path_kernel(7)


def test_batch_of_small_scenes(ct, O):
    """Several scenes, cameras, sizes and depths in shared per-path launches:
    each job's image equals the wavefront batch's and the job rendered alone.
    In the batch the glass scene's lanes (65 primitives) walk while the box's
    and the night scene's lanes scan; alone, those two run the scan-only
    build."""
    specs = [(scenes.cornell_box(), 40, 30, 3, 1, 8), (scenes.night_emissives(), 32, 24, 2, 7, 5),
             (scenes.glass_spheres(O.Rng(20240607)), 36, 24, 2, 3, 6)]
    jobs = {}
    for mode in (ct.PATHS_PER_PATH, ct.PATHS_WAVEFRONT):
        js = []
        for sc, w, h, spp, seed, depth in specs:
            ps = O.pack_scene(sc)
            js.append({"scene": ps, "frame": O.camera_frame(sc["camera"], w / h), "width": w,
                       "height": h, "spp": spp, "seed": seed, "maxBounces": depth,
                       "out": ct.DeviceBuffer.empty((w * h * 3,), np.float32)})
        tr = ct.Tracer()
        tr.set_path_kernel(mode)
        tr.render_batch(js)
        assert tr.get_path_kernel()[1] == (1 if mode == ct.PATHS_PER_PATH else 0)
        jobs[mode] = [j["out"].numpy() for j in js]
        if mode == ct.PATHS_PER_PATH:
            for j, m in zip(js, jobs[mode]):
                alone, _, _, _ = _render(ct, j["scene"], j["frame"], j, ct.PATHS_PER_PATH)
                assert np.array_equal(alone.view(np.uint32), m.view(np.uint32))
    for a, b in zip(jobs[ct.PATHS_PER_PATH], jobs[ct.PATHS_WAVEFRONT]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_graph_replays_and_counting(ct, O):
    sc, ps = _scene(ct, O, "cornell")
    fr = O.camera_frame(sc["camera"], 1.5)
    opts = {"width": 48, "height": 32, "spp": 4, "seed": 2, "maxBounces": 8}
    ref, sref, _, _ = _render(ct, ps, fr, opts, ct.PATHS_WAVEFRONT)
    tr = ct.Tracer()
    tr.set_path_kernel(ct.PATHS_PER_PATH)
    tr.set_graph(True)
    d = ct.DeviceBuffer.empty((48 * 32 * 3,), np.float32)
    for k in range(4):  # capture, then replays with eager work between them
        tr.render_image(ps, fr, opts, d)
        assert np.array_equal(d.numpy().view(np.uint32), ref.view(np.uint32)), k
        ct.renderImage(ps, fr, dict(opts, seed=k + 10))
    _, sc_, rc, _ = _render(ct, ps, fr, opts, ct.PATHS_PER_PATH, counting=True)
    assert rc == 1 and sc_["extensionRays"] == sref["extensionRays"]
    # the box (14 primitives) is scanned: every extension ray tests every
    # primitive, NEE rays stop at their first blocker, no node is visited
    # unless a tie sends a ray through the ordered walk
    tests = sc_["triTests"] + sc_["sphereTests"]
    assert tests >= sc_["extensionRays"] * ps.primCount
    assert 0 < sc_["shadowTriTests"] + sc_["shadowSphereTests"] <= sc_["shadowRays"] * ps.primCount
    assert sc_["shadowNodeVisits"] == 0
    # above the scan's limit (the mirror room, 2016 primitives) the rays walk
    sm, pm = _scene(ct, O, "mirror")
    fm = O.camera_frame(sm["camera"], 1.5)
    _, sm_, rm, _ = _render(ct, pm, fm, opts, ct.PATHS_PER_PATH, counting=True)
    assert rm == 1 and sm_["nodeVisits"] > 0 and sm_["shadowNodeVisits"] > 0


def test_radiance_per_path_equals_wavefront(ct, O):
    """radianceInto jobs (caller rays and RNG states) through both forms."""
    sc = ct.presetById("night")["scene"]
    ps = O.pack_scene(sc)
    rs = np.random.default_rng(4)
    n = 500
    o = rs.uniform(-1, 1, (n, 3))
    v = rs.normal(size=(n, 3))
    rays = np.concatenate([o, v / np.linalg.norm(v, axis=1, keepdims=True)], 1)
    states = rs.integers(1, 1 << 32, n).astype(np.uint32)
    out = {}
    for mode in (ct.PATHS_PER_PATH, ct.PATHS_WAVEFRONT):
        tr = ct.Tracer()
        tr.set_path_kernel(mode)
        d_rays = ct.DeviceBuffer.from_array(np.ascontiguousarray(rays))
        d_rng = ct.DeviceBuffer.from_array(states.copy())
        d_out = ct.DeviceBuffer.empty((n * 3,), np.float64)
        tr.radiance(ps, d_rays, n, d_rng, {"maxBounces": 8}, d_out)
        out[mode] = (d_out.numpy(), d_rng.numpy())
    (La, Sa), (Lb, Sb) = out[ct.PATHS_PER_PATH], out[ct.PATHS_WAVEFRONT]
    assert np.array_equal(La.view(np.uint64), Lb.view(np.uint64)) and np.array_equal(Sa, Sb)
    for i in range(0, n, 50):
        oL, ost = O.radiance(ps, rays[i, :3], rays[i, 3:], int(states[i]), max_bounces=8,
                             with_state=True)
        assert np.array_equal(La.reshape(-1, 3)[i], oL) and Sa[i] == ost


def test_scan_settles_ties_like_the_ordered_walk(ct, O):
    """Scenes of at most 64 primitives are scanned in packed order by the
    per-path launch, which on an exact-t tie meets the lower index first.
    Rays aimed at the Cornell box's shared edges and vertices, where triangles
    of different walls are hit at the same t, must shade the primitive
    intersectClosest keeps (intersect.ts:202-282): radiance and RNG states
    equal to the wavefront's (4-wide walk) and to the oracle's."""
    sc = ct.presetById("cornell")["scene"]
    ps = O.pack_scene(sc)
    assert ps.primCount <= 64
    rays = scenes.tie_rays(ps, 5, 12)
    ordered = O.closest_batch(ps, rays, walk=O.WALK_ORDERED)
    lowest = O.closest_batch(ps, rays, walk=O.WALK_STACKLESS)  # lowest index on ties
    tied = np.flatnonzero(ordered["prim"] != lowest["prim"])
    obj = ps.primObject
    cross = [i for i in tied if obj[ordered["prim"][i]] != obj[lowest["prim"][i]]]
    assert len(cross) >= 10  # ties between walls: the winner changes the shading
    n = len(rays)
    states = np.random.default_rng(6).integers(1, 1 << 32, n).astype(np.uint32)
    out = {}
    for mode in (ct.PATHS_PER_PATH, ct.PATHS_WAVEFRONT):
        tr = ct.Tracer()
        tr.set_path_kernel(mode)
        tr.set_counting(True)
        d_rays = ct.DeviceBuffer.from_array(np.ascontiguousarray(rays))
        d_rng = ct.DeviceBuffer.from_array(states.copy())
        d_out = ct.DeviceBuffer.empty((n * 3,), np.float64)
        tr.radiance(ps, d_rays, n, d_rng, {"maxBounces": 8}, d_out)
        out[mode] = (d_out.numpy().reshape(-1, 3), d_rng.numpy(), tr.stats(),
                     tr.get_path_kernel()[1])
    (La, Sa, sa, ra), (Lb, Sb, _, rb) = out[ct.PATHS_PER_PATH], out[ct.PATHS_WAVEFRONT]
    assert ra == 1 and rb == 0
    assert sa["nodeVisits"] > 0  # tied rays were walked again in the reference's order
    assert np.array_equal(La.view(np.uint64), Lb.view(np.uint64)) and np.array_equal(Sa, Sb)
    for i in cross + list(range(0, n, 97)):
        oL, ost = O.radiance(ps, rays[i, :3], rays[i, 3:], int(states[i]), max_bounces=8,
                             with_state=True)
        assert np.array_equal(La[i], oL) and Sa[i] == ost, i


@pytest.mark.parametrize("name", ["cornell", "mirror", "glass"])
def test_per_path_skip_link_walk(ct, O, name):
    """With the tracer's walk set to the skip-link walk the per-path launch
    walks with it too (no LDS stack): the same image again."""
    sc, ps = _scene(ct, O, name)
    fr = O.camera_frame(sc["camera"], 1.5)
    opts = {"width": 48, "height": 32, "spp": 3, "seed": 8, "maxBounces": 10}
    gs, ss, rs, _ = _render(ct, ps, fr, opts, ct.PATHS_PER_PATH, walk=ct.WALK_STACKLESS)
    gw, sw, _, _ = _render(ct, ps, fr, opts, ct.PATHS_WAVEFRONT)
    assert rs == 1
    assert np.array_equal(gs.view(np.uint32), gw.view(np.uint32))
    assert ss["extensionRays"] == sw["extensionRays"] and ss["shadowRays"] == sw["shadowRays"]


@pytest.mark.parametrize("graph", [False, True], ids=["eager", "graph"])
def test_mixed_batch_packs_each_form_in_its_own_rounds(ct, O, graph):
    """Small scenes and a 40k-triangle soup in one batch: under CT_PATHS_AUTO
    the small scenes' jobs run a per-path round before the soup's wavefront
    round (eager; a graph-mode tracer keeps one wavefront round), and forced
    per-path everything runs per-path.  Every job's image equals the
    wavefront batch's and the job rendered alone."""
    soup = ct.synth.soup_scene(40000, n_lights=16, n_materials=20, seed=9)
    specs = [(soup, 48, 32, 3, 2, 6), (scenes.cornell_box(), 40, 30, 3, 1, 8),
             (scenes.night_emissives(), 32, 24, 2, 7, 5)]
    packed = [O.pack_scene(sc) for sc, *_ in specs]
    out = {}
    for mode in (ct.PATHS_AUTO, ct.PATHS_PER_PATH, ct.PATHS_WAVEFRONT):
        js = [{"scene": ps, "frame": O.camera_frame(sc["camera"], w / h), "width": w, "height": h,
               "spp": spp, "seed": seed, "maxBounces": depth,
               "out": ct.DeviceBuffer.empty((w * h * 3,), np.float32)}
              for ps, (sc, w, h, spp, seed, depth) in zip(packed, specs)]
        tr = ct.Tracer()
        tr.set_path_kernel(mode)
        tr.set_graph(graph)
        for rep in range(3 if graph else 1):  # capture, then replays
            tr.render_batch(js)
            want = {ct.PATHS_PER_PATH: 1, ct.PATHS_WAVEFRONT: 0, ct.PATHS_AUTO: 0 if graph else 1}
            assert tr.get_path_kernel()[1] == want[mode]
            assert tr.stats()["rounds"] == (2 if mode == ct.PATHS_AUTO and not graph else 1)
            imgs = [j["out"].numpy() for j in js]
            if rep == 0:
                out[mode] = imgs
            else:
                for a, b in zip(imgs, out[mode]):
                    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), rep
        st = tr.stats()
        out[(mode, "rays")] = (st["extensionRays"], st["shadowRays"])
    for mode in (ct.PATHS_AUTO, ct.PATHS_PER_PATH):
        for a, b in zip(out[mode], out[ct.PATHS_WAVEFRONT]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert out[(mode, "rays")] == out[(ct.PATHS_WAVEFRONT, "rays")]
    for ps, (sc, w, h, spp, seed, depth), img in zip(packed, specs, out[ct.PATHS_PER_PATH]):
        opts = {"width": w, "height": h, "spp": spp, "seed": seed, "maxBounces": depth}
        alone, _, _, _ = _render(ct, ps, O.camera_frame(sc["camera"], w / h), opts,
                                 ct.PATHS_WAVEFRONT)
        assert np.array_equal(alone.view(np.uint32), img.view(np.uint32))
