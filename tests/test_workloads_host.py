"""Host-side workload generators and serving helpers (CPU only)."""
import math

import numpy as np


def test_batch_workload_shape(ct):
    counts = ct.synth.triangle_counts(64, 4, 1_000, 500_000)
    assert counts.min() >= 1_000 and counts.max() <= 500_000
    assert 10_000 <= np.median(counts) <= 40_000  # lognormal around 20k
    scenes, jobs = ct.synth.batch_workload(6, lo=1_000, hi=3_000)
    assert len(scenes) == len(jobs) == 6
    for i, (sc, j) in enumerate(zip(scenes, jobs)):
        assert j["sceneIndex"] == i and j["width"] == j["height"]
        assert j["width"] in (128, 256, 512, 768) and 1 <= j["spp"] <= 32
        assert j["maxBounces"] == 8
        tris = sum(len(o["shape"]["indices"]) // 3 for o in sc["objects"])
        assert tris == ct.synth.triangle_counts(6, 4, 1_000, 3_000)[i] + 64


def test_request_stream_is_poisson_with_stated_ranges(ct):
    reqs = ct.synth.request_stream(128, rate=200.0, duration=30.0, seed=11)
    n = len(reqs)
    assert abs(n - 6000) < 5 * math.sqrt(6000)
    at = np.array([r["at"] for r in reqs])
    assert np.all(np.diff(at) > 0) and at[-1] < 30.0
    gaps = np.diff(at)
    assert abs(gaps.mean() - 1 / 200) < 0.1 / 200  # exponential gaps, mean 1/rate
    assert abs(gaps.std() / gaps.mean() - 1) < 0.1
    for r in reqs:
        assert 0 <= r["sceneIndex"] < 128
        assert r["width"] == r["height"] and r["width"] % 128 == 0 and 128 <= r["width"] <= 1024
        assert 1 <= r["spp"] <= 64 and 4 <= r["maxBounces"] <= 12


def test_percentile_nearest_rank(ct):
    xs = list(range(1, 101))
    assert ct.serve.percentile(xs, 50) == 50
    assert ct.serve.percentile(xs, 99) == 99
    assert ct.serve.percentile(xs, 100) == 100
    assert ct.serve.percentile([7.0], 99) == 7.0


def test_cornell_c1_workload_is_the_test_variant(ct):
    # the small plain workload the bench times is the 256x256 Cornell variant the
    # parity tests render against the oracle
    import scenes

    assert ct.synth.cornell_c1() == scenes.cornell_c1()
    ref = ct.presets.cornell_box()
    c1 = {o["name"]: o["material"] for o in ct.synth.cornell_c1()["objects"]}
    assert c1["Ceiling light"]["strength"] == 15 and ref["objects"][5]["material"]["strength"] == 12
    assert c1["Metal sphere"]["roughness"] == 0.1
    assert c1["Back wall"]["checker"] == {"color2": (0.25, 0.25, 0.25), "scale": 8}
    assert "checker" not in ref["objects"][2]["material"]
