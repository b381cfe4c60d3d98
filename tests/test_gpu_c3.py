"""The precision form's workload on the GPU: the caustic scene (256 tessellated
glass spheres of 16,384 triangles over a 65,536-triangle checker plane, one
small bright light; 4,259,842 primitives) and the layout extension it needs.

From 2^21 primitives the float32 leaf words start*8+count of the reference
layout (bvh.ts:60-62) are no longer exact, so packScene also carries the exact
uint32 words, which every walk decodes leaves from.  The oracle reads the
same extension, so parity is checked at that size too:

* a 2.1M-triangle soup: GPU build = oracle build (nodes, order, exact words),
  closest hits of the binary64 walks = the oracle's, a render within 1e-4 per
  pixel;
* a 16-sphere caustic scene (327,682 primitives) in both precisions: the
  quantized FP32 walk returns the exhaustive FP32 scan's hits, the binary64
  render matches the oracle per pixel, the FP32 render tracks it;
* the full scene: a valid tree, closest hits equal to the oracle's over the
  same arrays, renders in both precisions within the same bounds.
"""
import numpy as np
import pytest

from conftest import oracle_packed

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def rel_err(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    den = np.maximum(np.abs(g), np.abs(o))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(den > 0, np.abs(g - o) / den, 0.0)


def rms(a, b):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)))


def _rays(ps, n, seed):
    """Rays from a box 1.4x the root box aimed at points inside it (binary64)."""
    nodes = np.asarray(ps.nodes, np.float32)
    lo, hi = nodes[0:3].astype(np.float64), nodes[4:7].astype(np.float64)
    c, h = (lo + hi) / 2, (hi - lo) / 2
    u = np.random.default_rng(seed).random((n, 6))
    o = c + (2 * u[:, 0:3] - 1) * h * 1.4
    d = c + (2 * u[:, 3:6] - 1) * h - o
    return np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], 1)


def _render(ct, ps, frame, opts, precision):
    tr = ct.Tracer()
    tr.set_precision(precision)
    out = ct.DeviceBuffer.empty((opts["width"] * opts["height"] * 3,), np.float32)
    tr.render_image(ps, frame, opts, out)
    return out.numpy()


@pytest.fixture(scope="module")
def soup2m(ct):
    sc = ct.synth.soup_scene(2_100_000, n_lights=16, n_materials=50, seed=21)
    print("2.1M soup: packing with the GPU builder", flush=True)
    return sc, ct.packScene(sc)


def test_exact_leaf_words_build_above_2_pow_21(ct, O, soup2m):
    sc, ps = soup2m
    assert ps.primCount >= 1 << 21 and ps.leafWords is not None
    kind, geom, _ = O.expand_scene(sc)
    bmin, bmax, bcen = O.prim_bounds(kind, geom)
    print("2.1M soup: oracle build", flush=True)
    on, onc, oo, ow = O.build_bvh(bmin, bmax, bcen, 12, leaf_words=True)
    assert ps.nodeCount == onc
    assert np.array_equal(ps.nodes.view(np.uint32), on.view(np.uint32))
    assert np.array_equal(ps.order, oo)
    assert np.array_equal(ps.leafWords, ow)
    # the float32 words really are inexact here, the exact ones decode to a valid tree
    leaf = ow != 0xFFFFFFFF
    assert (ps.nodes[7::8][leaf].astype(np.float64) != ow[leaf]).any()
    assert O.validate_bvh(on, onc, oo, ps.primCount, ow)["errors"] == 0


def test_walks_and_render_above_2_pow_21(ct, O, soup2m):
    sc, ps = soup2m
    view = oracle_packed(O, ps)
    rays = _rays(ps, 4096, 5)
    ref = O.closest_batch(view, rays, walk=O.WALK_ORDERED)
    g = ct.closest_hit(ps, rays, walk=ct.WALK_WIDE)
    assert (ref["prim"] >= 0).sum() > 1000
    assert np.array_equal(g["prim"], ref["prim"])
    assert np.array_equal(g["t"], ref["t"])
    # the binary walks decode leaves from the exact words too
    g = ct.closest_hit(ps, rays, walk=ct.WALK_ORDERED)
    assert np.array_equal(g["prim"], ref["prim"]) and np.array_equal(g["t"], ref["t"])
    sl = O.closest_batch(view, rays[:1024], walk=O.WALK_STACKLESS)
    g = ct.closest_hit(ps, rays[:1024], walk=ct.WALK_STACKLESS)
    assert np.array_equal(g["prim"], sl["prim"]) and np.array_equal(g["t"], sl["t"])
    w, h = 48, 32
    fr = ct.cameraFrame(sc["camera"], w / h)
    opts = {"width": w, "height": h, "spp": 2, "seed": 4, "maxBounces": 4}
    g64 = _render(ct, ps, fr, opts, ct.PRECISION_F64)
    o64 = O.render_image(view, fr, w, h, 2, seed=4, max_bounces=4)
    assert rel_err(g64, o64).max() <= REL_TOL
    g32 = _render(ct, ps, fr, opts, ct.PRECISION_F32)
    assert np.isfinite(g32).all()


@pytest.fixture(scope="module")
def caustic16(ct):
    sc = ct.synth.caustic_scene(n_spheres=16)
    return sc, ct.packScene(sc)


def test_reduced_caustic_scene_both_precisions(ct, O, caustic16):
    sc, ps = caustic16
    assert ps.primCount == 16 * 16384 + 65536 + 2
    r32 = _rays(ps, 20000, 9).astype(np.float32)
    pw, tw, _, _ = ct.closest_hit32(ps, r32, walk=ct.WALK_WIDE)
    pb, tb, _, _ = ct.closest_hit32(ps, r32, walk=ct.WALK_BRUTE)
    assert np.array_equal(tw.view(np.uint32), tb.view(np.uint32))
    assert (pw != pb).sum() <= 10  # exact ties in t only
    view = oracle_packed(O, ps)
    r64 = _rays(ps, 4096, 10)
    ref = O.closest_batch(view, r64, walk=O.WALK_ORDERED)
    g = ct.closest_hit(ps, r64, walk=ct.WALK_WIDE)
    assert np.array_equal(g["prim"], ref["prim"]) and np.array_equal(g["t"], ref["t"])
    w = h = 48
    fr = ct.cameraFrame(sc["camera"], 1.0)
    opts = {"width": w, "height": h, "spp": 8, "seed": 1, "maxBounces": 16}
    g64 = _render(ct, ps, fr, opts, ct.PRECISION_F64)
    o64 = O.render_image(view, fr, w, h, 8, seed=1, max_bounces=16)
    err = rel_err(g64, o64)
    print(f"caustic16 binary64 vs oracle: max rel err {err.max():.2e}")
    assert err.max() <= REL_TOL
    g32 = _render(ct, ps, fr, opts, ct.PRECISION_F32)
    g64b = _render(ct, ps, fr, dict(opts, seed=2), ct.PRECISION_F64)
    same, noise = rms(g32, g64), rms(g64, g64b)
    print(f"caustic16 RMS FP32 vs binary64 same seed {same:.4g}, binary64 seed 1 vs 2 {noise:.4g}")
    assert np.isfinite(g32).all() and same < 0.5 * noise


@pytest.fixture(scope="module")
def caustic_full(ct):
    sc = ct.synth.caustic_scene()
    print("caustic scene: packing 4,259,842 primitives with the GPU builder", flush=True)
    return sc, ct.packScene(sc)


def test_full_caustic_scene(ct, O, caustic_full):
    sc, ps = caustic_full
    assert ps.primCount == 4_259_842 and ps.leafWords is not None
    assert O.validate_bvh(ps.nodes, ps.nodeCount, ps.order, ps.primCount,
                          ps.leafWords)["errors"] == 0
    view = oracle_packed(O, ps)
    r64 = _rays(ps, 2048, 12)
    ref = O.closest_batch(view, r64, walk=O.WALK_ORDERED)
    g = ct.closest_hit(ps, r64, walk=ct.WALK_WIDE)
    assert np.array_equal(g["prim"], ref["prim"]) and np.array_equal(g["t"], ref["t"])
    r32 = r64[:512].astype(np.float32)
    pw, tw, _, _ = ct.closest_hit32(ps, r32, walk=ct.WALK_WIDE)
    pb, tb, _, _ = ct.closest_hit32(ps, r32, walk=ct.WALK_BRUTE)
    assert np.array_equal(tw.view(np.uint32), tb.view(np.uint32))
    w = h = 32
    fr = ct.cameraFrame(sc["camera"], 1.0)
    opts = {"width": w, "height": h, "spp": 4, "seed": 7, "maxBounces": 16}
    g64 = _render(ct, ps, fr, opts, ct.PRECISION_F64)
    print("caustic scene: oracle render of 32x32x4", flush=True)
    o64 = O.render_image(view, fr, w, h, 4, seed=7, max_bounces=16)
    assert rel_err(g64, o64).max() <= REL_TOL
    g32 = _render(ct, ps, fr, opts, ct.PRECISION_F32)
    assert np.isfinite(g32).all() and g32.max() > 0
