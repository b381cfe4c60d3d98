"""Serving form on the GPU: continuous batching over resident scenes.

Each job's result must be the same alone and in the mix -- bit for bit, since
every pixel's samples are added in pass order whatever rounds they fall in --
and must match the oracle's renderImage within the 1e-4 per-pixel tolerance."""
import time

import numpy as np
import pytest

import scenes
from conftest import oracle_packed

pytestmark = pytest.mark.gpu


def _requests(ct):
    """Jobs of three scenes with different cameras, resolutions, spp and depths."""
    base = [ct.packScene(scenes.cornell_box()), ct.packScene(scenes.night_emissives()),
            ct.packScene(ct.synth.family_scene(3, 4000))]
    cams = [scenes.cornell_box()["camera"], scenes.night_emissives()["camera"],
            ct.synth.family_scene(3, 10)["camera"]]
    spec = [(0, 64, 48, 5, 8, 1), (1, 40, 30, 3, 5, 7), (2, 48, 48, 4, 6, 2), (0, 32, 32, 7, 4, 3),
            (2, 24, 24, 9, 12, 4), (1, 56, 40, 2, 8, 5), (0, 16, 16, 12, 8, 6)]
    out = []
    for si, w, h, spp, depth, seed in spec:
        out.append({"scene": base[si], "frame": ct.cameraFrame(cams[si], w / h), "width": w,
                    "height": h, "spp": spp, "maxBounces": depth, "seed": seed})
    return out


def _drained(srv, timeout=5.0):
    """True once the worker has freed every released buffer."""
    t = time.time()
    while srv.stats()["hbmInUse"] and time.time() - t < timeout:
        time.sleep(0.005)
    return srv.stats()["hbmInUse"] == 0


def _alone(ct, r):
    return ct.renderImage(r["scene"], r["frame"], r)


def rel_err(g, o):
    g = np.asarray(g, np.float64)
    o = np.asarray(o, np.float64)
    den = np.maximum(np.abs(g), np.abs(o))
    return np.where(den > 0, np.abs(g - o) / np.where(den > 0, den, 1), 0.0)


@pytest.mark.parametrize("opts", [
    {},                                            # one round takes everything
    {"max_paths": 4096, "slice_passes": 1},         # many rounds, jobs interleaved pass by pass
    {"max_paths": 8192, "policy": "srpt"},
    {"max_paths": 8192, "policy": "scene"},          # each scene's jobs together
    {"hbm_budget": 64 * 48 * 12 + 1, "max_paths": 16384},  # HBM admits about one job at a time
    {"max_active": 2, "max_paths": 8192, "graph": True}])
def test_alone_equals_in_mix(ct, O, opts):
    reqs = _requests(ct)
    with ct.Server(**opts) as srv:
        ids = []
        for k, r in enumerate(reqs):
            ids.append(srv.submit(r["scene"], r["frame"], r["width"], r["height"], r["spp"],
                                  seed=r["seed"], maxBounces=r["maxBounces"]))
            if k == 3:
                time.sleep(0.01)  # later arrivals join a running mix
        imgs = []
        for jid, r in zip(ids, reqs):
            st = srv.wait(jid, timeout=120)
            assert st["state"] == ct.serve.DONE and st["passesDone"] == r["spp"]
            assert st["doneSec"] >= st["admitSec"] >= st["submitSec"] >= 0
            imgs.append(srv.read(jid, r["width"], r["height"]))
            srv.release(jid)
        assert _drained(srv)
        stats = srv.stats()
        assert stats["completed"] == len(reqs)
        assert stats["pixelSamples"] == sum(r["width"] * r["height"] * r["spp"] for r in reqs)
        if "max_paths" in opts:
            assert stats["rounds"] > 1
    for r, img in zip(reqs, imgs):
        assert np.array_equal(_alone(ct, r).view(np.uint32), img.view(np.uint32))
    r = reqs[1]
    o = O.render_image(oracle_packed(O, r["scene"]), r["frame"], r["width"], r["height"], r["spp"], seed=r["seed"],
                       max_bounces=r["maxBounces"])
    assert rel_err(imgs[1], o).max() <= 1e-4


def test_progressive_read_and_cancel(ct):
    reqs = _requests(ct)
    r = reqs[0]
    with ct.Server(max_paths=4096, slice_passes=1) as srv:
        big = srv.submit(r["scene"], r["frame"], r["width"], r["height"], 40, seed=1)
        queued = [srv.submit(r["scene"], r["frame"], 16, 16, 2, seed=s) for s in range(3)]
        srv.release(queued[1])  # cancel (queued or running)
        st = srv.wait(big, timeout=120)
        assert st["passesDone"] == 40
        img = srv.read(big, r["width"], r["height"])
        assert np.isfinite(img).all()
        for q in (queued[0], queued[2]):
            assert srv.wait(q, timeout=120)["state"] == ct.serve.DONE
            srv.release(q)
        srv.release(big)
        with pytest.raises(ct.CTError):
            srv.poll(big)
        assert _drained(srv)


def test_submit_rejects_what_cannot_run(ct):
    r = _requests(ct)[0]
    with ct.Server(max_paths=4096, hbm_budget=1 << 20) as srv:
        with pytest.raises(ct.CTError):  # one pass larger than a round
            srv.submit(r["scene"], r["frame"], 128, 128, 1)
        with pytest.raises(ct.CTError):  # zero spp
            srv.submit(r["scene"], r["frame"], 16, 16, 0)
        with pytest.raises(ct.CTError):  # buffer larger than the HBM budget
            srv2 = ct.Server(hbm_budget=1000)
            try:
                srv2.submit(r["scene"], r["frame"], 32, 32, 1)
            finally:
                srv2.close()


def test_replay_stream_latencies(ct):
    """A short timed stream through `replay`: every request completes, and its
    latency is at least its own trace time and finite."""
    reqs = _requests(ct)
    stream = []
    for k in range(12):
        r = reqs[k % len(reqs)]
        stream.append({"at": 0.002 * k, "sceneIndex": k % len(reqs), "width": r["width"],
                       "height": r["height"], "spp": r["spp"], "maxBounces": r["maxBounces"],
                       "seed": 100 + k})
    scenes_ = [r["scene"] for r in reqs]
    frames = [reqs[k % len(reqs)]["frame"] for k in range(len(stream))]
    with ct.Server(max_paths=1 << 16) as srv:
        res = ct.serve.replay(srv, scenes_, frames, stream, keep={0, 5})
    assert all(x is not None and x > 0 for x in res["latency"])
    assert res["stats"]["completed"] == len(stream)
    for idx, img in res["images"].items():
        s = stream[idx]
        r = dict(reqs[idx % len(reqs)], seed=s["seed"])
        assert np.array_equal(_alone(ct, r).view(np.uint32), img.view(np.uint32))


def test_alone_equals_in_mix_fp32(ct):
    """Precision form through the server: the same image as the FP32 tracer
    renders the job alone (compensation terms persist across rounds)."""
    reqs = _requests(ct)
    with ct.Server(max_paths=4096, slice_passes=2, precision=ct.PRECISION_F32) as srv:
        ids = [srv.submit(r["scene"], r["frame"], r["width"], r["height"], r["spp"], seed=r["seed"],
                          maxBounces=r["maxBounces"]) for r in reqs]
        imgs = []
        for jid, r in zip(ids, reqs):
            srv.wait(jid, timeout=120)
            imgs.append(srv.read(jid, r["width"], r["height"]))
            srv.release(jid)
        assert srv.stats()["rounds"] > 1
    tr = ct.Tracer()
    tr.set_precision(ct.PRECISION_F32)
    for r, img in zip(reqs, imgs):
        out = ct.DeviceBuffer.empty((r["width"] * r["height"] * 3,), np.float32)
        tr.render_image(r["scene"], r["frame"], r, out)
        assert np.array_equal(out.numpy().view(np.uint32), img.view(np.uint32))


def test_read_of_a_job_queued_behind_a_round_filling_job(ct):
    """A job admitted while another one takes every path of the round gets its
    buffer allocated and cleared by the worker outside the lock; a read in that
    window must wait for the clear.  Every read returns the sum of the passes
    added so far (zeros before the first), never uninitialised memory."""
    r = _requests(ct)[0]
    w, h, spp, seed = 32, 24, 3, 11
    opts = {"width": w, "height": h, "spp": spp, "seed": seed, "maxBounces": 6}
    prefix = [np.zeros(w * h * 3, np.float32)]
    acc = np.zeros(w * h * 3, np.float32)
    for s in range(spp - 1):
        ct.accumulatePass(r["scene"], r["frame"], w, h, s + 7919 * seed, opts, acc)
        prefix.append(acc.copy())
    final = ct.renderImage(r["scene"], r["frame"], opts)
    valid = [p.view(np.uint32) for p in prefix] + [final.view(np.uint32)]
    # one pass of the big job is the whole round: the small job waits admitted
    with ct.Server(max_paths=128 * 128) as srv:
        big = srv.submit(r["scene"], r["frame"], 128, 128, 24, seed=1)
        small = srv.submit(r["scene"], r["frame"], w, h, spp, seed=seed, maxBounces=6)
        t0 = time.time()
        while srv.poll(small)["state"] == ct.serve.QUEUED and time.time() - t0 < 60:
            time.sleep(0.0002)
        seen = set()
        for _ in range(200):
            img = srv.read(small, w, h).view(np.uint32)
            k = [i for i, v in enumerate(valid) if np.array_equal(v, img)]
            assert k, "a read returned neither a pass sum nor the final image"
            seen.add(k[0])
            if srv.poll(small)["state"] == ct.serve.DONE:
                break
        assert srv.wait(small, timeout=120)["state"] == ct.serve.DONE
        img = srv.read(small, w, h)
        assert np.array_equal(img.view(np.uint32), valid[-1])
        assert srv.wait(big, timeout=120)["state"] == ct.serve.DONE
        srv.release(small)
        srv.release(big)
        assert _drained(srv)
