"""bench.py's host-side parts on the CPU: the CPU baseline (the JS restatement
on worker threads, the C oracle beside it, the C oracle alone when the JS leg
fails), its band sizing, and the serving form's fixed-cost fit."""
import numpy as np
import pytest

import bench


@pytest.fixture(scope="module")
def small(ct, O):
    sc = ct.synth.soup_scene(3000, n_lights=4, n_materials=8, seed=4)
    ps = O.pack_scene(sc)
    w, h = 64, 36
    return ps, O.camera_frame(sc["camera"], w / h), w, h


def _oracle_band(O, ps, frame, w, h, spp, depth, seed, rows):
    acc = np.zeros(w * h * 3, np.float32)
    for s in range(spp):
        O.accumulate_pass(ps, frame, w, h, s + 7919 * seed, acc, max_bounces=depth, rows=rows)
    return (acc[rows[0] * w * 3:rows[1] * w * 3].astype(np.float64) * (1.0 / spp)).astype(
        np.float32)


@pytest.mark.skipif(__import__("shutil").which("node") is None, reason="no node executable")
def test_cpu_baseline_js_leg(ct, O, small):
    ps, frame, w, h = small
    cpu, (rows, band) = bench.cpu_baseline(ct, ps, frame, w, h, 2, 3, 5, budget_s=0.3)
    assert cpu["impl"].startswith("JS restatement")
    assert cpu["cores"] == bench.host_cpus()["allowed"] and cpu["value"] > 0
    assert len(cpu["trials"]) == 3 and cpu["single_thread"]["value"] > 0
    assert cpu["c_oracle"]["value"] > 0
    assert 0 <= rows[0] < rows[1] <= h
    assert np.array_equal(band, _oracle_band(O, ps, frame, w, h, 2, 3, 5, rows))


def test_cpu_baseline_c_oracle_when_the_js_leg_fails(ct, O, small, monkeypatch):
    ps, frame, w, h = small

    def fail(*a, **k):
        raise RuntimeError("no worker threads here")

    monkeypatch.setattr(O, "render_band_js", fail)
    cpu, (rows, band) = bench.cpu_baseline(ct, ps, frame, w, h, 2, 3, 5, budget_s=0.3)
    assert cpu["impl"].startswith("C oracle") and "c_oracle" not in cpu
    assert cpu["value"] > 0 and 0 <= rows[0] < rows[1] <= h
    assert np.array_equal(band, _oracle_band(O, ps, frame, w, h, 2, 3, 5, rows))


def test_band_size():
    assert bench._band_size(0.5, 4, 3.0, 3, 100) == 8     # 2 sets of 4 rows per second
    assert bench._band_size(10.0, 4, 3.0, 3, 100) == 4    # at least one set
    assert bench._band_size(1e-9, 4, 3.0, 3, 100) == 100  # never beyond the frame


def test_fixed_cost_fit():
    reqs = [{"width": s, "height": s, "spp": p} for s, p in
            [(128, 1), (256, 4), (1024, 64), (512, 8), (128, 2), (768, 16)]]
    rays = np.array([3e4, 5e5, 1.3e8, 4e6, 6e4, 2.5e7])
    t = 0.0004 + rays * 0.6e-9
    f = bench._fixed_cost(reqs, t, rays, float(t.sum()))
    assert f["fixed_ms"] == pytest.approx(0.4, abs=1e-6)
    assert f["ns_per_ray"] == pytest.approx(0.6, abs=1e-6)
    assert f["small"]["requests"] == 3  # under 2^20 samples: 128^2x1, 256^2x4, 128^2x2
