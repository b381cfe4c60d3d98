"""Serving form: a request loop over device-resident scenes.

`Server` wraps the library's serving loop (continuous batching on its own
thread, progressive buffers resident in HBM, admission by HBM budget);
`replay` feeds it a timed request stream the way clients would and measures
time-to-complete per request.  Each job computes renderImage (tracer.ts:348)
of its scene and camera; the result is bit-identical to rendering the job
alone.
"""
import ctypes as C
import time

import numpy as np

from . import _capi as A
from ._capi import DeviceBuffer, check, lib
from .tracer import Tracer, device_scene

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3
POLICIES = {"fifo": 0, "srpt": 1, "scene": 2}


def _status(s):
    return {k: getattr(s, k) for k, _ in A.JobStatus._fields_}


class Server:
    def __init__(self, hbm_budget=0, max_paths=0, max_active=0, slice_passes=0, policy="fifo",
                 graph=False, precision=0):
        A.init()
        o = A.ServerOpts(int(hbm_budget), int(max_paths), int(max_active), int(slice_passes),
                         POLICIES[policy], int(bool(graph)), int(precision))
        h = C.c_void_p()
        check(lib().ct_server_create(C.byref(o), C.byref(h)), "ct_server_create")
        self.handle = h
        self._scenes = {}

    def submit(self, scene, frame, width, height, spp, seed=0, maxBounces=8,
               russianRoulette=True, nee=True):
        """Queues renderImage(scene, frame, {width, height, spp, seed, ...});
        returns the job id."""
        ds = device_scene(scene)
        self._scenes[id(ds)] = ds  # keep resident while jobs may use it
        j = A.Job()
        j.scene = ds.handle.value
        j.frame = Tracer._frame16(frame)
        j.width, j.height, j.spp, j.seed = width, height, spp, seed & 0xFFFFFFFF
        j.maxBounces, j.russianRoulette, j.nee = int(maxBounces), int(bool(russianRoulette)), \
            int(bool(nee))
        jid = C.c_uint64()
        check(lib().ct_job_submit(self.handle, C.byref(j), C.byref(jid)), "ct_job_submit")
        return jid.value

    def poll(self, jid):
        s = A.JobStatus()
        check(lib().ct_job_poll(self.handle, jid, C.byref(s)), "ct_job_poll")
        return _status(s)

    def wait(self, jid, timeout=None):
        s = A.JobStatus()
        check(lib().ct_job_wait(self.handle, jid, -1.0 if timeout is None else float(timeout),
                                C.byref(s)), "ct_job_wait")
        return _status(s)

    def read(self, jid, width, height, out=None):
        """The job's buffer: the final image once done (host float32 array, or
        copied into the DeviceBuffer `out`)."""
        if isinstance(out, DeviceBuffer):
            check(lib().ct_job_read(self.handle, jid, out.ptr, 1), "ct_job_read")
            return out
        a = np.empty(width * height * 3, np.float32)
        check(lib().ct_job_read(self.handle, jid, a.ctypes.data_as(C.c_void_p), 0), "ct_job_read")
        return a

    def read_rgba(self, jid, width, height, exposureEv=0.0, mode="aces"):
        """The job's frame tone mapped on the device (tonemap.ts:41-62): the final
        image once done, the progressive average before. RGBA8 numpy array."""
        a = np.empty(width * height * 4, np.uint8)
        check(lib().ct_job_read_rgba(self.handle, jid, float(exposureEv),
                                     {"aces": 0, "reinhard": 1}[mode],
                                     a.ctypes.data_as(C.c_void_p), 0), "ct_job_read_rgba")
        return a

    def release(self, jid):
        check(lib().ct_job_release(self.handle, jid), "ct_job_release")

    def stats(self):
        s = A.ServerStats()
        check(lib().ct_server_stats_get(self.handle, C.byref(s)), "ct_server_stats_get")
        return {k: getattr(s, k) for k, _ in A.ServerStats._fields_}

    def close(self):
        if self.handle is not None and self.handle.value:
            check(lib().ct_server_destroy(self.handle), "ct_server_destroy")
        self.handle = None
        self._scenes.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def percentile(xs, q):
    """Nearest-rank percentile (q in [0, 100]) of a non-empty sequence."""
    xs = sorted(xs)
    k = max(0, min(len(xs) - 1, int(np.ceil(q / 100.0 * len(xs))) - 1))
    return xs[k]


def replay(server, scenes, frames, requests, keep=None, poll_s=0.0005, log=None):
    """Submits each request at its arrival time `at` (seconds from the start),
    collects completions as they happen and releases their buffers.

    scenes[i] / frames[k]: the resident scene of request k's sceneIndex and its
    camera frame.  keep: optional set of request indices whose final images are
    read back before release (returned in `images`).  Returns per-request
    latency (done - submit, server clock), the wall time of the whole replay,
    the server's counters and the largest delay between a request's arrival
    time and its submission (`submit_lag_max`, seconds)."""
    pending = {}
    lat = [None] * len(requests)
    images = {}
    t0 = time.perf_counter()
    k = 0
    last_log = t0
    lag = 0.0
    while k < len(requests) or pending:
        now = time.perf_counter() - t0
        while k < len(requests) and requests[k]["at"] <= now:
            r = requests[k]
            lag = max(lag, now - r["at"])
            jid = server.submit(scenes[r["sceneIndex"]], frames[k], r["width"], r["height"],
                                r["spp"], seed=r["seed"], maxBounces=r["maxBounces"],
                                russianRoulette=r.get("russianRoulette", True),
                                nee=r.get("nee", True))
            pending[jid] = k
            k += 1
        done = []
        for jid, idx in pending.items():
            s = server.poll(jid)
            if s["state"] == FAILED:
                raise RuntimeError(f"request {idx} failed: {lib().ct_last_error().decode()}")
            if s["state"] == DONE:
                lat[idx] = s["doneSec"] - s["submitSec"]
                if keep is not None and idx in keep:
                    r = requests[idx]
                    images[idx] = server.read(jid, r["width"], r["height"])
                done.append(jid)
        for jid in done:
            server.release(jid)
            del pending[jid]
        if log is not None and time.perf_counter() - last_log > 10:
            last_log = time.perf_counter()
            st = server.stats()
            log(f"replay t={now:.1f}s submitted {k}/{len(requests)} completed {st['completed']} "
                f"queued {st['queued']} active {st['active']}")
        if not done:
            time.sleep(poll_s)
    wall = time.perf_counter() - t0
    return {"latency": lat, "wall": wall, "stats": server.stats(), "images": images,
            "submit_lag_max": lag}
