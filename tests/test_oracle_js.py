"""The JS restatement of the reference tracer (oracle/js/trace.js), which the
CPU baseline runs on Node worker threads, against the C oracle: the same
float32 sums bit for bit and the same extension and NEE ray counts, over
scenes that reach every branch -- triangle and sphere lights (outside and
inside the sphere), metal, dielectric, checker, the thin lens, the sky
gradient, Russian roulette and NEE off, the exact-leaf-word layout."""
import numpy as np
import pytest

import scenes

node_missing = pytest.mark.skipif(
    __import__("shutil").which("node") is None, reason="no node executable")


def _check(O, ps, frame, w, h, spp, seed, depth, rows, threads=2, **kw):
    rep, sums = O.render_band_js(ps, frame, w, h, spp, seed, rows, threads, max_bounces=depth,
                                 **kw)
    acc = np.zeros(w * h * 3, np.float32)
    ext = sh = 0
    for s in range(spp):
        c = O.accumulate_pass(ps, frame, w, h, s + 7919 * seed, acc, max_bounces=depth,
                              rows=rows, **kw)
        ext += int(c[0])
        sh += int(c[1])
    assert rep["threads"] == threads and len(rep["trials"]) == 1
    tr = rep["trials"][0]
    assert (tr["extension"], tr["shadow"]) == (ext, sh)
    assert np.array_equal(acc.view(np.uint32), sums.view(np.uint32))
    assert ext > 0
    return acc


@node_missing
@pytest.mark.parametrize("name", ["cornell", "cornell_c1", "night", "glass", "dof", "mirror"])
def test_js_tracer_equals_oracle(ct, O, name):
    if name == "cornell":
        sc = scenes.cornell_box()
    elif name == "cornell_c1":
        sc = scenes.cornell_c1()
    else:
        sc = ct.presetById(name)["scene"]
    ps = O.pack_scene(sc)
    w, h = 40, 30
    fr = O.camera_frame(sc["camera"], w / h)
    acc = _check(O, ps, fr, w, h, 3, 5, 8, (6, 20), threads=3)
    assert np.isfinite(acc).all() and acc.max() > 0


@node_missing
def test_js_tracer_options_and_sky(ct, O):
    sc = ct.presetById("glass")["scene"]
    assert sc["sky"]["kind"] == "gradient"
    ps = O.pack_scene(sc)
    w, h = 32, 24
    fr = O.camera_frame(sc["camera"], w / h)
    _check(O, ps, fr, w, h, 2, 1, 12, (0, h), threads=1, russian_roulette=False)
    _check(O, ps, fr, w, h, 2, 2, 5, (3, 9), threads=2, nee=False)


@node_missing
def test_js_tracer_exact_leaf_words(ct, O):
    # a soup whose oracle view carries the exact-word layout extension (the
    # words equal the float32 ones at this size, so both readings must agree)
    sc = ct.synth.soup_scene(6000, n_lights=8, n_materials=12, seed=3)
    ps = O.pack_scene(sc)
    f = ps.nodes[7::8]
    words = np.where(f >= 0, f, -1).astype(np.int64).astype(np.uint32)
    view = O.Packed(ps.prims, ps.nodes, ps.nodeCount, ps.mats, ps.lights, ps.primObject,
                    ps.order, ps.sky, words)
    w, h = 32, 20
    fr = O.camera_frame(sc["camera"], w / h)
    a = _check(O, view, fr, w, h, 2, 4, 6, (5, 15), threads=2)
    b = _check(O, ps, fr, w, h, 2, 4, 6, (5, 15), threads=2)
    assert np.array_equal(a, b)


@node_missing
def test_js_tracer_inside_a_sphere_light(O):
    # a diffuse floor inside an emissive sphere: NEE takes the inside branch
    # of sampleLightInto (tracer.ts:138-153), a sphere light from outside the
    # cone branch -- both in one scene
    from test_oracle_pins import CAM, _emitter, _lambert, _quad, _sphere
    ps = O.pack_scene({"objects": [
        {"name": "floor", "shape": _quad((-30.0, 0.0, -30.0), (60.0, 0.0, 0.0),
                                         (0.0, 0.0, 60.0)), "material": _lambert(0.5)},
        {"name": "dome", "shape": _sphere((0.0, 0.0, 0.0), 10.0), "material": _emitter(1.0)},
        {"name": "lamp", "shape": _sphere((1.0, 2.0, 0.5), 0.3), "material": _emitter(4.0)}],
        "camera": CAM, "sky": {"kind": "none"}})
    w = h = 16
    frame = O.camera_frame(dict(CAM, distance=5.0, fovDeg=30.0), 1.0)
    acc = _check(O, ps, frame, w, h, 4, 9, 3, (0, h), threads=2)
    assert acc.max() > 0
