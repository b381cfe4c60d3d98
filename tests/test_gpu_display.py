"""Display and picking on the GPU: tonemapBuffer (tonemap.ts:41-62) against the
oracle's restatement, the reference's golden PPM through a GPU render and the
GPU tone map, served frames as RGBA, and pick (picking.ts:21-33) with the
check of scenes.test.ts:94-104."""
import os

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cornell_96x64_16spp.ppm")


@pytest.mark.parametrize("mode,ev,samples", [("aces", 0.0, 1), ("aces", 1.5, 7),
                                             ("reinhard", -0.5, 3), ("aces", 0.0, 0)])
def test_tonemap_matches_oracle(ct, O, mode, ev, samples):
    rng = np.random.default_rng(3)
    w, h = 173, 61
    acc = (rng.lognormal(-1.0, 1.5, w * h * 3) * samples).astype(np.float32)
    acc[::97] = 0
    acc[5::101] = -0.25  # negative sums clamp to black
    acc[7::103] = 1e6
    g = ct.tonemapBuffer(acc, w, h, samples, ev, mode)
    o = O.tonemap_buffer(acc, w, h, samples, ev, mode)
    d = np.abs(g.astype(np.int32) - o.astype(np.int32))
    assert d.max() <= 1 and (d > 0).sum() <= 2  # at most a last-ulp pow() rounding flip
    assert (g.reshape(-1, 4)[:, 3] == 255).all()


def test_golden_ppm_through_gpu_render_and_tonemap(ct, O):
    sc = scenes.cornell_box()
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 96 / 64)
    img = ct.renderImage(ps, fr, {"width": 96, "height": 64, "spp": 16, "seed": 1,
                                  "maxBounces": 8, "russianRoulette": True, "nee": True})
    rgba = ct.tonemapBuffer(img, 96, 64, 1, 0.0, "aces")
    ppm = ct.encodePPM(96, 64, rgba)
    gold = open(GOLDEN, "rb").read()
    assert ppm[:20] == gold[:20]
    w, h, a = ct.decodePPM(ppm)
    _, _, b = ct.decodePPM(gold)
    assert (w, h) == (96, 64)
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    assert d.max() <= 1 and (d > 0).mean() < 0.002


def test_served_frames_as_rgba(ct):
    sc = scenes.cornell_box()
    ps = ct.packScene(sc)
    fr = ct.cameraFrame(sc["camera"], 1.0)
    with ct.Server(max_paths=4096, slice_passes=1) as srv:
        jid = srv.submit(ps, fr, 32, 32, 6, seed=2)
        st = srv.wait(jid, timeout=120)
        assert st["state"] == ct.serve.DONE
        final = srv.read(jid, 32, 32)
        rgba = srv.read_rgba(jid, 32, 32, 0.5, "aces")
        assert np.array_equal(rgba, ct.tonemapBuffer(final, 32, 32, 1, 0.5, "aces"))
        srv.release(jid)


def test_pick_cornell_glass_sphere(ct):
    """scenes.test.ts:94-104: the glass sphere lies lower right of centre."""
    sc = ct.presets.CORNELL_BOX
    ps = ct.packScene(sc)
    frame = ct.cameraFrame(sc["camera"], 1.5)
    hit = ct.pick(ps, frame, 0.74, 0.15)
    assert hit is not None
    assert sc["objects"][hit["objectIndex"]]["name"] == "Glass sphere"
    assert hit["t"] > 1
    o, d = ct.primaryRay(frame, 0.74, 0.15)
    assert np.allclose(hit["point"], np.array(o) + hit["t"] * np.array(d))
    empty = ct.packScene(dict(sc, objects=[]))
    assert ct.pick(empty, frame, 0.5, 0.5) is None
