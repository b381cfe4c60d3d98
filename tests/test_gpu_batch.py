"""Multi-job launches (batching and serving forms) on the GPU.

Every job of a shared launch must produce exactly (bit for bit) what it
produces alone, and what the oracle's renderImage / accumulatePass produce
within the 1e-4 per-pixel tolerance; graph replays must equal eager runs."""
import numpy as np
import pytest

import scenes
from conftest import oracle_packed

pytestmark = pytest.mark.gpu


def _jobs(ct, O):
    sc1 = scenes.cornell_box()
    sc2 = scenes.night_emissives()
    sc3 = ct.synth.soup_scene(40000, n_lights=16, n_materials=20, seed=9)
    out = []
    for sc, (w, h, spp, seed, depth) in ((sc1, (64, 48, 4, 1, 8)), (sc2, (40, 30, 2, 7, 5)),
                                         (sc3, (48, 32, 3, 2, 6))):
        ps = O.pack_scene(sc)
        out.append({"scene": ps, "frame": O.camera_frame(sc["camera"], w / h), "width": w,
                    "height": h, "spp": spp, "seed": seed, "maxBounces": depth})
    return out


def _alloc(ct, j):
    return ct.DeviceBuffer.empty((j["width"] * j["height"] * 3,), np.float32)


def rel_err(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    den = np.maximum(np.abs(g), np.abs(o))
    return np.where(den > 0, np.abs(g - o) / np.where(den > 0, den, 1), 0.0)


def test_batch_equals_alone_and_oracle(ct, O):
    jobs = _jobs(ct, O)
    tr = ct.Tracer()
    for j in jobs:
        j["out"] = _alloc(ct, j)
    tr.render_batch(jobs)
    st = tr.stats()
    mixed = [j["out"].numpy() for j in jobs]
    for j, m in zip(jobs, mixed):
        alone = ct.renderImage(j["scene"], j["frame"], j)
        assert np.array_equal(alone.view(np.uint32), m.view(np.uint32))
        o = O.render_image(j["scene"], j["frame"], j["width"], j["height"], j["spp"],
                           seed=j["seed"], max_bounces=j["maxBounces"])
        assert rel_err(m, o).max() <= 1e-4
    # the two small scenes' per-path round, then the soup's wavefront round
    assert st["rounds"] == 2


def test_batch_split_into_rounds(ct, O):
    jobs = _jobs(ct, O)
    tr = ct.Tracer()
    tr.set_capacity(4096)  # forces many rounds, one pass of each job at most per round
    for j in jobs:
        j["out"] = _alloc(ct, j)
    tr.render_batch(jobs)
    assert tr.stats()["rounds"] > 3
    for j in jobs:
        alone = ct.renderImage(j["scene"], j["frame"], j)
        assert np.array_equal(alone.view(np.uint32), j["out"].numpy().view(np.uint32))


def test_progressive_passes_resume(ct, O):
    """(accum, passes done) resumes a render: 1 + 2 passes equal 3 in one call,
    and equal the oracle's accumulatePass sums."""
    jobs = _jobs(ct, O)
    tr = ct.Tracer()
    a = [_alloc(ct, j) for j in jobs]
    b = [_alloc(ct, j) for j in jobs]
    for buf in a + b:
        buf.upload(np.zeros(buf.shape, np.float32))

    def pj(j, acc, first, passes):
        return {**j, "frameIndex": first, "passes": passes, "accum": acc}

    tr.accumulate_batch([pj(j, x, 100, 1) for j, x in zip(jobs, a)])
    tr.accumulate_batch([pj(j, x, 101, 2) for j, x in zip(jobs, a)])
    tr.accumulate_batch([pj(j, x, 100, 3) for j, x in zip(jobs, b)])
    for j, x, y in zip(jobs, a, b):
        assert np.array_equal(x.numpy().view(np.uint32), y.numpy().view(np.uint32))
        acc = np.zeros(j["width"] * j["height"] * 3, np.float32)
        for f in (100, 101, 102):
            O.accumulate_pass(j["scene"], j["frame"], j["width"], j["height"], f, acc,
                              max_bounces=j["maxBounces"])
        assert rel_err(x.numpy(), acc).max() <= 1e-4


def test_graph_replays_equal_eager(ct, O):
    jobs = _jobs(ct, O)
    eager = ct.Tracer()
    for j in jobs:
        j["out"] = _alloc(ct, j)
    eager.render_batch(jobs)
    ref = [j["out"].numpy() for j in jobs]
    tr = ct.Tracer()
    tr.set_graph(True)
    for rep in range(3):
        for j in jobs:
            j["out"].upload(np.full(j["out"].shape, np.nan, np.float32))
        tr.render_batch(jobs)
        ct.sync()
        got = [j["out"].numpy() for j in jobs]
        for r, g in zip(ref, got):
            assert np.array_equal(r.view(np.uint32), g.view(np.uint32)), rep
        # eager work between replays
        eager.render_batch(jobs[:1])


def _family_jobs(ct):
    """Five small scenes of the batching family (soups from seeds 0..4) with
    uneven sizes, resolutions and sample counts."""
    out = []
    for i, (n, side, spp) in enumerate(((1000, 32, 1), (1500, 48, 2), (3000, 24, 3),
                                        (1200, 40, 2), (2500, 16, 4))):
        sc = ct.synth.family_scene(i, n)
        out.append({"scene": ct.packScene(sc), "frame": ct.cameraFrame(sc["camera"], 1.0),
                    "width": side, "height": side, "spp": spp, "seed": i + 1, "maxBounces": 8})
    for j in out:
        j["out"] = _alloc(ct, j)
    return out


def _images(jobs):
    return [j["out"].numpy().view(np.uint32) for j in jobs]


@pytest.mark.parametrize("setting", [
    {"workgroup": 64}, {"workgroup": 128}, {"walk": 1, "shadowWalk": 1, "stackDepth": 8},
    {"walk": 0, "shadowWalk": 0}, {"walk": 2, "shadowWalk": 0, "stackDepth": 16},
    {"persistent": 0}, {"maxPaths": 4096}, {"blocks": 64, "workgroup": 64},
    {"pathGroups": 1}, {"pathGroups": 2, "maxPaths": 1 << 21}])
def test_launch_settings_do_not_change_images(ct, O, setting):
    jobs = _family_jobs(ct)
    ref_tr = ct.Tracer()
    ref_tr.render_batch(jobs)
    ref = _images(jobs)
    tr = ct.Tracer()
    cfg = tr.get_config()
    cfg.update(setting)
    tr.set_config(cfg)
    assert tr.get_config() == cfg
    for j in jobs:
        j["out"].upload(np.full(j["out"].shape, np.nan, np.float32))
    tr.render_batch(jobs)
    for r, g in zip(ref, _images(jobs)):
        assert np.array_equal(r, g)
    # and the oracle agrees with one of them
    j = jobs[2]
    o = O.render_image(oracle_packed(O, j["scene"]), j["frame"], j["width"], j["height"], j["spp"], seed=j["seed"],
                       max_bounces=j["maxBounces"])
    assert rel_err(j["out"].numpy(), o).max() <= 1e-4


def test_set_config_is_all_or_nothing(ct):
    tr = ct.Tracer()
    before = tr.get_config()
    bad = dict(before, workgroup=96, walk=0)
    with pytest.raises(ct.CTError):
        tr.set_config(bad)
    assert tr.get_config() == before


def test_traversal_grid_follows_precision_until_set(ct):
    """The persistent traversal launches default to 640 workgroups in binary64
    and 768 in FP32; an explicit schedule sticks across precision changes."""
    tr = ct.Tracer()
    cfg = tr.get_config()
    assert cfg["persistent"] == 1 and cfg["blocks"] == 640
    tr.set_precision(ct.PRECISION_F32)
    assert tr.get_config()["blocks"] == 768
    tr.set_schedule(True, 512)
    tr.set_precision(ct.PRECISION_F64)
    assert tr.get_config()["blocks"] == 512
    tr.set_schedule(False, 1)
    assert tr.get_config()["persistent"] == 0
    tr.set_config(dict(tr.get_config(), persistent=1, blocks=700))
    assert tr.get_config()["blocks"] == 700


def test_autotune_picks_a_setting_and_caches_it(ct, tmp_path):
    jobs = _family_jobs(ct)
    ct.Tracer().render_batch(jobs)
    ref = _images(jobs)
    cache = tmp_path / "tune.txt"
    tr = ct.Tracer()
    res = tr.autotune(jobs, max_trials=6, cache_path=str(cache))
    assert not res["fromCache"] and 1 <= res["trials"] <= 6
    assert res["ms"] > 0 and res["ms"] <= res["initialMs"]
    assert tr.get_config() == res["config"]
    for r, g in zip(ref, _images(jobs)):
        assert np.array_equal(r, g)
    assert len(cache.read_text().splitlines()) == 1
    for j in jobs:
        j["out"].upload(np.zeros(j["out"].shape, np.float32))
    tr2 = ct.Tracer()
    res2 = tr2.autotune(jobs, max_trials=6, cache_path=str(cache))
    assert res2["fromCache"] and res2["trials"] == 0 and res2["config"] == res["config"]
    assert tr2.get_config() == res["config"]
    for r, g in zip(ref, _images(jobs)):
        assert np.array_equal(r, g)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_presets_in_one_batch(ct, O, precision):
    """All five presets of scenes.ts in one render_batch (spheres, the mesh,
    lens cameras and scenes without a sky share the multi-scene launches):
    each image equals the same job rendered alone, and in binary64 the
    oracle's within 1e-4."""
    sizes = [(48, 32, 3, 5, 8), (40, 40, 2, 6, 6), (32, 24, 4, 7, 8), (56, 32, 2, 8, 4),
             (40, 30, 3, 9, 12)]
    jobs = []
    for p, (w, h, spp, seed, depth) in zip(ct.presets.SCENE_PRESETS, sizes):
        sc = p["scene"]
        jobs.append({"scene": O.pack_scene(sc), "frame": O.camera_frame(sc["camera"], w / h),
                     "width": w, "height": h, "spp": spp, "seed": seed, "maxBounces": depth})
    tr = ct.Tracer()
    tr.set_precision(ct.PRECISION_F32 if precision == "f32" else ct.PRECISION_F64)
    for j in jobs:
        j["out"] = _alloc(ct, j)
    tr.render_batch(jobs)
    assert tr.stats()["rounds"] == 1
    alone_tr = ct.Tracer()
    alone_tr.set_precision(ct.PRECISION_F32 if precision == "f32" else ct.PRECISION_F64)
    for j in jobs:
        m = j["out"].numpy()
        d = _alloc(ct, j)
        alone_tr.render_image(j["scene"], j["frame"], j, d)
        assert np.array_equal(d.numpy().view(np.uint32), m.view(np.uint32))
        assert np.isfinite(m).all() and m.max() > 0
        if precision == "f64":
            o = O.render_image(j["scene"], j["frame"], j["width"], j["height"], j["spp"],
                               seed=j["seed"], max_bounces=j["maxBounces"])
            assert rel_err(m, o).max() <= 1e-4
