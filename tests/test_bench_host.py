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


def test_other_forms_summaries_and_failures(monkeypatch):
    """The default line's "forms" object: each form's summary keeps its own
    figures, the shorter settings reach the forms, and a form that fails is
    recorded with its error while the others still report."""
    import argparse

    seen = {}

    def fake_batch(ct, a):
        seen["batch"] = (a.steps, a.warmup, a.no_cpu)
        return {"config": {"workload": "batch w"}, "value": 1.0, "unit": "Mrays/s",
                "dtype": "f64", "ms_per_step": 2.0, "spp_mpix_per_s": 3.0, "rounds": 1,
                "plain_sequential": {"ms": 4.0}, "speedup_vs_plain": 1.5,
                "identical_to_plain": True, "extra": "dropped",
                "autotune": {"trials": 2, "ms": 2.0, "config": {}, "initialMs": 3.0},
                "roofline": {"kernel": "k", "achieved": 1.0, "frac": 0.1, "peak": 8000.0}}

    def fake_serve(ct, a):
        seen["serve"] = (a.duration, a.rates)
        raise RuntimeError("no GPU here")

    monkeypatch.setattr(bench, "batch_form", fake_batch)
    monkeypatch.setattr(bench, "serve_form", fake_serve)
    args = argparse.Namespace(forms=["batch", "serve", "bogus"], steps=5, warmup=1, no_cpu=False,
                              forms_serve_duration=5.0, duration=2.0, rates=[20.0])
    out = bench.other_forms(None, args)
    assert seen["batch"] == (3, 1, True) and seen["serve"] == (5.0, [])
    b = out["batch"]
    assert b["workload"] == "batch w" and b["speedup_vs_plain"] == 1.5 and "extra" not in b
    assert b["autotune"] == {"trials": 2, "ms": 2.0, "config": {}}
    assert b["roofline"] == {"kernel": "k", "achieved": 1.0, "frac": 0.1}
    assert out["serve"]["error"] == "RuntimeError: no GPU here"
    assert out["bogus"]["error"].startswith("ValueError")
    assert all("run_s" in v for v in out.values())
    assert args.steps == 5 and args.duration == 2.0  # the caller's settings are untouched
