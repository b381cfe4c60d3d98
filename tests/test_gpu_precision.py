"""Precision form on the GPU: quantized 4-wide nodes, FP32 arithmetic,
compensated accumulation.

* The quantized walk is conservative: its closest hits are those of the
  exhaustive FP32 scan, bit for bit (t always; the primitive too except on
  exact ties in t), and its occlusion answers are the scan's.
* The FP32 primitive tests are those of a binary32 numpy restatement.
* Images follow the binary64 form's distribution: the same RNG stream drives
  both, so an FP32 image differs from the FP64 one far less than two FP64
  images of different seeds do; the furnace energy balance (reduced in FP64
  on the GPU) holds.
"""
import numpy as np
import pytest

import fp32_ref
import scenes
from conftest import oracle_packed

pytestmark = pytest.mark.gpu


def _scene(ct, O, name):
    if name == "cornell":
        return scenes.cornell_box()
    if name == "soup":
        return ct.synth.soup_scene(20000, n_lights=16, n_materials=20, seed=3)
    if name == "random":
        return scenes.random_scene(O.Rng(1234), 1000, 2000)
    if name == "coincident":
        return scenes.coincident_spheres()
    if name == "glass":
        return scenes.glass_spheres(O.Rng(20240607))
    raise KeyError(name)


def _rays32(ct, ps, n, seed):
    """Rays from a box 1.4x the scene's (root node box) aimed at points of the
    scene box, in the shape of bvh.test.ts:94-99; float32."""
    nodes = np.asarray(ps.nodes, np.float32)
    lo = nodes[0:3].astype(np.float64)
    hi = nodes[4:7].astype(np.float64)
    c, h = (lo + hi) / 2, (hi - lo) / 2
    g = np.random.default_rng(seed)
    u = g.random((n, 6))
    o = c + (2 * u[:, 0:3] - 1) * h * 1.4
    t = c + (2 * u[:, 3:6] - 1) * h
    v = t - o
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return np.concatenate([o, v], 1).astype(np.float32)


@pytest.mark.parametrize("name", ["cornell", "soup", "random", "coincident", "glass"])
def test_quantized_walk_equals_exhaustive_scan(ct, O, name):
    ps = ct.packScene(_scene(ct, O, name))
    rays = _rays32(ct, ps, 20000, 11)
    pw, tw, uw, vw = ct.closest_hit32(ps, rays, walk=ct.WALK_WIDE)
    pb, tb, ub, vb = ct.closest_hit32(ps, rays, walk=ct.WALK_BRUTE)
    assert np.array_equal(tw.view(np.uint32), tb.view(np.uint32))
    diff = pw != pb
    assert diff.sum() <= max(2, len(rays) // 2000)  # exact ties in t only
    # u, v on triangle hits (a sphere hit leaves them as the last triangle
    # test set them, intersect.ts:173-176, which depends on the visiting order)
    kind = np.asarray(ps.prims, np.float32).reshape(-1, 16)[:, 3]
    tri = ~diff & (pw >= 0) & (kind[np.maximum(pw, 0)] != 0)
    assert np.array_equal(uw[tri].view(np.uint32), ub[tri].view(np.uint32))
    assert np.array_equal(vw[tri].view(np.uint32), vb[tri].view(np.uint32))
    assert (pw >= 0).sum() > len(rays) // 20
    tmax = np.random.default_rng(3).uniform(0.5, 40, len(rays)).astype(np.float32)
    ow = ct.occluded32(ps, rays, tmax, walk=ct.WALK_WIDE)
    ob = ct.occluded32(ps, rays, tmax, walk=ct.WALK_BRUTE)
    assert np.array_equal(ow, ob)
    c = ct.closest_hit32_counts(ps, rays)
    assert c["nodes"] > 0 and c["tris"] + c["spheres"] > 0


@pytest.mark.parametrize("name", ["cornell", "random"])
def test_fp32_scan_matches_binary32_restatement(ct, O, name):
    ps = ct.packScene(_scene(ct, O, name))
    rays = _rays32(ct, ps, 1500, 21)
    p, t, _, _ = ct.closest_hit32(ps, rays, walk=ct.WALK_BRUTE)
    pr, tr = fp32_ref.brute_closest(np.asarray(ps.prims, np.float32)[: ps.primCount * 16], rays)
    assert np.array_equal(p, pr)
    assert np.array_equal(t.view(np.uint32), tr.view(np.uint32))


def _render(ct, ps, frame, opts, precision, stack=24, workgroup=256):
    tr = ct.Tracer()
    tr.set_precision(precision)
    tr.set_traversal(ct.WALK_WIDE, stack)
    tr.set_workgroup(workgroup)
    out = ct.DeviceBuffer.empty((opts["width"] * opts["height"] * 3,), np.float32)
    tr.render_image(ps, frame, opts, out)
    return out.numpy(), tr


def test_precision_render_tracks_binary64(ct, O):
    sc = scenes.cornell_c1()
    ps = ct.packScene(sc)
    fr = ct.cameraFrame(sc["camera"], 1.0)
    opts = {"width": 64, "height": 64, "spp": 32, "seed": 1, "maxBounces": 8}
    i32, tr = _render(ct, ps, fr, opts, ct.PRECISION_F32)
    i64, _ = _render(ct, ps, fr, opts, ct.PRECISION_F64)
    i64b, _ = _render(ct, ps, fr, dict(opts, seed=2), ct.PRECISION_F64)
    assert np.isfinite(i32).all() and i32.max() > 0
    rms = lambda a, b: float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)))
    assert rms(i32, i64) < 0.3 * rms(i64, i64b)
    m32 = i32.reshape(-1, 3).mean(0)
    m64 = i64.reshape(-1, 3).mean(0)
    assert np.all(np.abs(m32 - m64) / m64 < 0.01)
    st = tr.stats()
    assert st["extensionRays"] > 0 and st["shadowRays"] > 0
    # the FP64 image of this form is the oracle's (per-pixel parity holds there)
    o = O.render_image(oracle_packed(O, ps), fr, 64, 64, 32, seed=1, max_bounces=8)
    den = np.maximum(np.abs(i64), np.abs(o)).astype(np.float64)
    rel = np.where(den > 0, np.abs(i64.astype(np.float64) - o) / np.where(den > 0, den, 1), 0)
    assert rel.max() <= 1e-4


def test_precision_settings_and_batching_do_not_change_images(ct):
    sc = ct.synth.soup_scene(20000, n_lights=16, n_materials=20, seed=3)
    ps = ct.packScene(sc)
    fr = ct.cameraFrame(sc["camera"], 1.0)
    opts = {"width": 48, "height": 48, "spp": 6, "seed": 3, "maxBounces": 6}
    ref, _ = _render(ct, ps, fr, opts, ct.PRECISION_F32)
    for stack, wg in ((8, 256), (16, 64), (24, 128)):
        img, _ = _render(ct, ps, fr, opts, ct.PRECISION_F32, stack, wg)
        assert np.array_equal(ref.view(np.uint32), img.view(np.uint32)), (stack, wg)
    # batched with other jobs and split into rounds: the same image
    other = ct.packScene(scenes.cornell_box())
    tr = ct.Tracer()
    tr.set_precision(ct.PRECISION_F32)
    tr.set_capacity(4096)
    jobs = [{"scene": ps, "frame": fr, **opts},
            {"scene": other, "frame": ct.cameraFrame(scenes.cornell_box()["camera"], 1.0),
             "width": 40, "height": 30, "spp": 3, "seed": 2, "maxBounces": 5}]
    for j in jobs:
        j["out"] = ct.DeviceBuffer.empty((j["width"] * j["height"] * 3,), np.float32)
    tr.render_batch(jobs)
    assert tr.stats()["rounds"] > 2
    assert np.array_equal(ref.view(np.uint32), jobs[0]["out"].numpy().view(np.uint32))


def test_compensated_accumulation_matches_render(ct):
    """accumulatePass in FP32 folds the call's compensation into the sum, so
    renderImage's result is exactly that sum times 1/spp (same passes)."""
    sc = scenes.cornell_box()
    ps = ct.packScene(sc)
    fr = ct.cameraFrame(sc["camera"], 1.0)
    tr = ct.Tracer()
    tr.set_precision(ct.PRECISION_F32)
    a = ct.DeviceBuffer.from_array(np.zeros(32 * 32 * 3, np.float32))
    tr.accumulate_batch([{"scene": ps, "frame": fr, "width": 32, "height": 32, "frameIndex": 0,
                          "passes": 3, "maxBounces": 8, "accum": a}])
    img = ct.DeviceBuffer.empty((32 * 32 * 3,), np.float32)
    tr.render_image(ps, fr, {"width": 32, "height": 32, "spp": 3, "seed": 0, "maxBounces": 8}, img)
    acc = a.numpy()
    assert np.isfinite(acc).all() and acc.sum() > 0
    want = (acc * np.float32(1.0 / 3)).astype(np.float32)
    assert np.array_equal(want.view(np.uint32), img.numpy().view(np.uint32))


@pytest.mark.parametrize("nee,rr,seed", [(False, False, 3), (True, False, 9), (False, True, 5)])
def test_furnace_fp32_with_fp64_reduction(ct, nee, rr, seed):
    sc = scenes.furnace()
    ps = ct.packScene(sc)
    fr = ct.cameraFrame(sc["camera"], 1.0)
    w = 8
    tr = ct.Tracer()
    tr.set_precision(ct.PRECISION_F32)
    out = ct.DeviceBuffer.empty((w * w * 3,), np.float32)
    tr.render_image(ps, fr, {"width": w, "height": w, "spp": 4096, "seed": seed,
                             "maxBounces": 64, "russianRoulette": rr, "nee": nee}, out)
    sums = ct.image_sum_f64(out, w * w)
    host = out.numpy().astype(np.float64).reshape(-1, 3).sum(0)
    assert np.allclose(sums, host, rtol=1e-12)
    mean = sums.sum() / (3 * w * w)
    assert abs(mean - 5) / 5 < 0.05
