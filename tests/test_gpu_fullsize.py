"""The plain form's configurations at their stated sizes, against the oracle.

* Cornell variant: 256x256, 16 spp, maxBounces 8, RR and NEE, 16-bin SAH
  BVH -- every pixel within 1e-4 of the oracle's render.
* 1M-triangle soup (+64 emissive triangles): the GPU-built BVH bit-identical
  to the oracle's build, and the 1920x1080, 4 spp, maxBounces 6 frame of the
  throughput workload within 1e-4 per pixel over a band of rows (the oracle
  renders those rows with the same per-pixel seeds).
"""
import os

import numpy as np
import pytest

import scenes
from conftest import oracle_packed

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


def rel_err(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    den = np.maximum(np.abs(g), np.abs(o))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(den > 0, np.abs(g - o) / den, 0.0)


def _threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 4


def _oracle_rows(O, view, frame, w, h, spp, seed, depth, rows):
    """renderImage of rows [y0, y1) by the oracle: spp passes with frame index
    s + 7919*seed, then times 1/spp, in float32 as tracer.ts:348-355."""
    acc = np.zeros(w * h * 3, np.float32)
    for s in range(spp):
        O.accumulate_pass(view, frame, w, h, s + 7919 * seed, acc, max_bounces=depth,
                          threads=_threads(), rows=rows)
    return (acc.astype(np.float64) * (1.0 / spp)).astype(np.float32)


def test_cornell_variant_full_size(ct, O):
    sc = scenes.cornell_c1()
    ps = ct.packScene(sc, bins=16)
    ops = O.pack_scene(sc, bins=16)
    assert np.array_equal(ps.nodes.view(np.uint32), ops.nodes.view(np.uint32))
    w = h = 256
    fr = ct.cameraFrame(sc["camera"], 1.0)
    g = ct.renderImage(ps, fr, {"width": w, "height": h, "spp": 16, "seed": 1, "maxBounces": 8})
    print("cornell 256x256x16: oracle render", flush=True)
    o = _oracle_rows(O, ops, fr, w, h, 16, 1, 8, (0, h))
    err = rel_err(g, o)
    print(f"cornell 256x256x16: max rel err {err.max():.2e}, exact {(g == o).mean():.4f}")
    assert np.isfinite(g).all() and err.max() <= REL_TOL


@pytest.fixture(scope="module")
def soup1m(ct):
    sc = ct.synth.soup_scene()
    print("1M soup: packing with the GPU builder", flush=True)
    return sc, ct.packScene(sc)


def test_soup_1m_bvh_equals_oracle_build(ct, O, soup1m):
    sc, ps = soup1m
    kind, geom, _ = O.expand_scene(sc)
    bmin, bmax, bcen = O.prim_bounds(kind, geom)
    print("1M soup: oracle build", flush=True)
    on, onc, oo = O.build_bvh(bmin, bmax, bcen, 12)
    assert ps.primCount == 1_000_064 and ps.nodeCount == onc
    assert np.array_equal(ps.nodes.view(np.uint32), on.view(np.uint32))
    assert np.array_equal(ps.order, oo)


def test_soup_1m_1080p_frame_band(ct, O, soup1m):
    sc, ps = soup1m
    w, h, spp, depth, seed = 1920, 1080, 4, 6, 0
    fr = ct.cameraFrame(sc["camera"], w / h)
    g = ct.renderImage(ps, fr, {"width": w, "height": h, "spp": spp, "seed": seed,
                                "maxBounces": depth})
    rows = (528, 552)
    print(f"1M soup 1080p: oracle rows {rows}", flush=True)
    o = _oracle_rows(O, oracle_packed(O, ps), fr, w, h, spp, seed, depth, rows)
    band = slice(rows[0] * w * 3, rows[1] * w * 3)
    err = rel_err(g[band], o[band])
    print(f"1M soup 1080p rows {rows}: max rel err {err.max():.2e}, "
          f"exact {(g[band] == o[band]).mean():.4f}")
    assert np.isfinite(g).all() and g[band].max() > 0
    assert err.max() <= REL_TOL
