"""Host mirror of the reference's render/intersect API over the HIP library.

Names and argument meaning follow src/tracer/tracer.ts, intersect.ts and camera.ts;
the work runs on the GPU through the C ABI.  A scene is any object with the
PackedScene fields of pack.ts:46-59 (prims, primCount, nodes, nodeCount, mats,
matCount, lights, primObject, sky) -- the arrays the reference's packer emits.
"""
import ctypes as C
import math
import os

import numpy as np

from . import _capi as A
from ._capi import DeviceBuffer, check, lib

T_MIN = 1e-4  # intersect.ts:14
WALK_STACKLESS = 0
WALK_ORDERED = 1
WALK_WIDE = 2
WALK_BRUTE = 3
PRECISION_F64 = 0
PRECISION_F32 = 1
PATHS_AUTO = 0  # per-path launches for rounds over small scenes, else the wavefront
PATHS_WAVEFRONT = 1
PATHS_PER_PATH = 2


class DeviceScene:
    """Device-resident copy of a PackedScene (replaces handing it to the CPU tracer)."""

    def __init__(self, ps):
        A.init()
        sky = A.Sky()
        sk = getattr(ps, "sky", None) or {"kind": "none"}
        if sk.get("kind") == "gradient":
            sky.kind = 1
            sky.zenith = (C.c_double * 3)(*sk["zenith"])
            sky.horizon = (C.c_double * 3)(*sk["horizon"])
            sky.strength = sk["strength"]
        prims = np.ascontiguousarray(ps.prims, np.float32)
        nodes = np.ascontiguousarray(ps.nodes, np.float32)
        mats = np.ascontiguousarray(ps.mats, np.float32)
        lights = np.ascontiguousarray(ps.lights, np.int32)
        pobj = np.ascontiguousarray(ps.primObject, np.uint32)
        h = C.c_void_p()
        args = (prims.ctypes.data_as(C.c_void_p), ps.primCount, nodes.ctypes.data_as(C.c_void_p),
                ps.nodeCount, mats.ctypes.data_as(C.c_void_p), ps.matCount,
                lights.ctypes.data_as(C.c_void_p), len(lights), pobj.ctypes.data_as(C.c_void_p),
                C.byref(sky))
        words = getattr(ps, "leafWords", None)
        if words is not None:
            words = np.ascontiguousarray(words, np.uint32)
            check(lib().ct_scene_create_ex(*args, words.ctypes.data_as(C.c_void_p), C.byref(h)),
                  "ct_scene_create_ex")
        else:
            check(lib().ct_scene_create(*args, C.byref(h)), "ct_scene_create")
        self.handle = h
        self.primCount = ps.primCount
        self.nodeCount = ps.nodeCount

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().ct_scene_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_scene(ps):
    """The device copy of `ps`, uploaded once and cached on the object."""
    if isinstance(ps, DeviceScene):
        return ps
    ds = getattr(ps, "_device_scene", None)
    if ds is None:
        ds = DeviceScene(ps)
        try:
            ps._device_scene = ds
        except AttributeError:
            pass
    return ds


def _rays_on_device(rays):
    if isinstance(rays, DeviceBuffer):
        return rays, rays.shape[0] if len(rays.shape) > 1 else rays.nbytes // 48
    r = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
    return DeviceBuffer.from_array(r), len(r)


def closest_hit(ps, rays, tmax=None, walk=WALK_ORDERED, tmin=T_MIN, to_host=True):
    """Batched intersectClosest (walk=WALK_ORDERED, intersect.ts:202) or
    intersectClosestStackless (WALK_STACKLESS, intersect.ts:290).

    Returns dict prim/t/u/v (numpy when to_host, else DeviceBuffers)."""
    ds = device_scene(ps)
    d_rays, n = _rays_on_device(rays)
    d_tmax = None
    if tmax is not None:
        d_tmax = tmax if isinstance(tmax, DeviceBuffer) else DeviceBuffer.from_array(
            np.ascontiguousarray(np.broadcast_to(np.asarray(tmax, np.float64), (n,))))
    out = {"prim": DeviceBuffer.empty((n,), np.int32), "t": DeviceBuffer.empty((n,), np.float64),
           "u": DeviceBuffer.empty((n,), np.float64), "v": DeviceBuffer.empty((n,), np.float64)}
    check(lib().ct_closest_hit(ds.handle, d_rays.ptr, n, tmin, d_tmax.ptr if d_tmax else None,
                               walk, out["prim"].ptr, out["t"].ptr, out["u"].ptr, out["v"].ptr,
                               None), "ct_closest_hit")
    if to_host:
        return {k: v.numpy() for k, v in out.items()}
    return out


def closest_hit_counts(ps, rays, tmax=None, walk=WALK_ORDERED, tmin=T_MIN):
    """(node fetches, triangle tests, sphere tests) of a closest-hit batch."""
    ds = device_scene(ps)
    d_rays, n = _rays_on_device(rays)
    d_tmax = None
    if tmax is not None:
        d_tmax = DeviceBuffer.from_array(
            np.ascontiguousarray(np.broadcast_to(np.asarray(tmax, np.float64), (n,))))
    c = (C.c_uint64 * 3)()
    check(lib().ct_closest_hit_counts(ds.handle, d_rays.ptr, n, tmin,
                                      d_tmax.ptr if d_tmax else None, walk, c, None),
          "ct_closest_hit_counts")
    return int(c[0]), int(c[1]), int(c[2])


def occluded(ps, rays, tmax, tmin=T_MIN):
    """Batched occluded (intersect.ts:368): bool per ray."""
    ds = device_scene(ps)
    d_rays, n = _rays_on_device(rays)
    d_tmax = DeviceBuffer.from_array(
        np.ascontiguousarray(np.broadcast_to(np.asarray(tmax, np.float64), (n,))))
    out = DeviceBuffer.empty((n,), np.uint8)
    check(lib().ct_occluded(ds.handle, d_rays.ptr, n, tmin, d_tmax.ptr, out.ptr, None),
          "ct_occluded")
    return out.numpy().astype(bool)


def _rays32_on_device(rays):
    if isinstance(rays, DeviceBuffer):
        return rays, rays.nbytes // 24
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    return DeviceBuffer.from_array(r), r.shape[0]


def closest_hit32(ps, rays, tmax=None, walk=WALK_WIDE, tmin=T_MIN):
    """Precision form: FP32 closest hit over the quantized 4-wide nodes
    (walk=WALK_WIDE) or every primitive (WALK_BRUTE). rays: (n, 6) float32.
    Returns host arrays prim, t, u, v (float32)."""
    ds = device_scene(ps)
    d_rays, n = _rays32_on_device(rays)
    d_tmax = None if tmax is None else DeviceBuffer.from_array(np.ascontiguousarray(tmax, np.float32))
    outs = [DeviceBuffer.empty((max(1, n),), t) for t in (np.int32, np.float32, np.float32, np.float32)]
    check(lib().ct_closest_hit32(ds.handle, d_rays.ptr, n, C.c_float(tmin),
                                 None if d_tmax is None else d_tmax.ptr, walk,
                                 *[o.ptr for o in outs], None), "ct_closest_hit32")
    return tuple(o.numpy()[:n] for o in outs)


def closest_hit32_counts(ps, rays, tmax=None, tmin=T_MIN):
    """Work of the quantized walk: {nodes (4 per node), tris, spheres, fallbacks}."""
    ds = device_scene(ps)
    d_rays, n = _rays32_on_device(rays)
    d_tmax = None if tmax is None else DeviceBuffer.from_array(np.ascontiguousarray(tmax, np.float32))
    c = DeviceBuffer.from_array(np.zeros(4, np.uint64))
    check(lib().ct_closest_hit32_counts(ds.handle, d_rays.ptr, n, C.c_float(tmin),
                                        None if d_tmax is None else d_tmax.ptr, c.ptr, None),
          "ct_closest_hit32_counts")
    h = c.numpy()
    return {"nodes": int(h[0]), "tris": int(h[1]), "spheres": int(h[2]), "fallbacks": int(h[3])}


def occluded32(ps, rays, tmax, walk=WALK_WIDE, tmin=T_MIN):
    ds = device_scene(ps)
    d_rays, n = _rays32_on_device(rays)
    d_tmax = DeviceBuffer.from_array(np.ascontiguousarray(tmax, np.float32))
    out = DeviceBuffer.empty((max(1, n),), np.uint8)
    check(lib().ct_occluded32(ds.handle, d_rays.ptr, n, C.c_float(tmin), d_tmax.ptr, walk,
                              out.ptr, None), "ct_occluded32")
    return out.numpy()[:n].astype(bool)


def image_sum_f64(img, npix):
    """Per-channel binary64 sums of a device float32 RGB image, reduced on the GPU."""
    out = np.zeros(3, np.float64)
    check(lib().ct_image_sum_f64(img.ptr, int(npix), out.ctypes.data_as(C.c_void_p), None),
          "ct_image_sum_f64")
    return out


def intersectClosest(ps, ox, oy, oz, dx, dy, dz, tMin, tMax, hit):
    """Single-ray form of intersect.ts:202 filling a HitRecord-like dict."""
    r = closest_hit(ps, np.array([[ox, oy, oz, dx, dy, dz]]), None if tMax == math.inf else tMax,
                    WALK_ORDERED, tMin)
    hit["t"], hit["prim"], hit["u"], hit["v"] = float(r["t"][0]), int(r["prim"][0]), \
        float(r["u"][0]), float(r["v"][0])
    return hit["prim"] >= 0


# ---------------------------------------------------------------------------
# camera.ts

def _normalize(v):
    ln = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    if ln > 0:
        s = 1 / ln
        return (v[0] * s, v[1] * s, v[2] * s)
    return (0.0, 0.0, 0.0)


def cameraFrame(cam, aspect):
    """cameraFrame (camera.ts:33-51) as the 16-double frame the C ABI takes."""
    t = cam["target"]
    cp = math.cos(cam["pitch"])
    d = cam["distance"]
    o = (t[0] + cp * math.sin(cam["yaw"]) * d, t[1] + math.sin(cam["pitch"]) * d,
         t[2] + cp * math.cos(cam["yaw"]) * d)
    f = _normalize((t[0] - o[0], t[1] - o[1], t[2] - o[2]))
    r = (f[1] * 0 - f[2] * 1, f[2] * 0 - f[0] * 0, f[0] * 1 - f[1] * 0)
    if r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < 1e-12:
        r = (1.0, 0.0, 0.0)
    r = _normalize(r)
    u = (r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0])
    fd = cam["focusDistance"]
    return np.array([*o, *r, *u, *f, math.tan((cam["fovDeg"] * math.pi) / 360), aspect,
                     cam["lensRadius"], fd if fd > 1e-3 else 1e-3], np.float64)


def primaryRay(frame, sx, sy):
    """primaryRay (camera.ts:75-77) = generateRay(frame, sx, sy, 0, 0)
    (camera.ts:62-72) in binary64: the pinhole ray through (sx, sy), y up."""
    f = [float(x) for x in frame]
    o, r, u, fw = f[0:3], f[3:6], f[6:9], f[9:12]
    th, aspect, lens, fd = f[12], f[13], f[14], f[15]
    px = (2 * sx - 1) * th * aspect
    py = (2 * sy - 1) * th
    focus = [o[k] + ((r[k] * px + u[k] * py) + fw[k]) * fd for k in range(3)]
    offset = [r[k] * (0.0 * lens) + u[k] * (0.0 * lens) for k in range(3)]
    origin = [o[k] + offset[k] for k in range(3)]
    d = _normalize([focus[k] - origin[k] for k in range(3)])
    return origin, d


def pick(ps, frame, sx, sy):
    """pick (picking.ts:21-33): the object under normalised screen position
    (sx, sy) in [0, 1] (sy = 1 at the top), found with the ordered walk on the
    GPU.  Returns {objectIndex, prim, t, point} or None on a miss."""
    o, d = primaryRay(frame, sx, sy)
    ray = np.array([[*o, *d]], np.float64)
    h = closest_hit(ps, ray, walk=WALK_ORDERED)
    prim = int(h["prim"][0])
    if prim < 0:
        return None
    t = float(h["t"][0])
    obj = np.asarray(ps.primObject, np.uint32)[prim] if not isinstance(ps, DeviceScene) else None
    return {"objectIndex": None if obj is None else int(obj), "prim": prim, "t": t,
            "point": (o[0] + d[0] * t, o[1] + d[1] * t, o[2] + d[2] * t)}


# ---------------------------------------------------------------------------
# tracer.ts

class Tracer:
    """Owns the wavefront path state on the device (one per concurrent caller)."""

    def __init__(self, walk=WALK_WIDE):
        A.init()
        h = C.c_void_p()
        check(lib().ct_tracer_create(C.byref(h)), "ct_tracer_create")
        self.handle = h
        self.set_walk(walk)

    def set_walk(self, walk):
        check(lib().ct_tracer_set_walk(self.handle, walk), "ct_tracer_set_walk")

    def set_traversal(self, shadow_walk=WALK_WIDE, stack=24):
        check(lib().ct_tracer_set_traversal(self.handle, shadow_walk, stack), "set_traversal")

    def set_schedule(self, persistent=True, blocks=None):
        """Persistent traversal launches of `blocks` workgroups (default: the
        tracer's current count), or one lane per path when not persistent."""
        if blocks is None:
            blocks = self.get_config()["blocks"]
        check(lib().ct_tracer_set_schedule(self.handle, int(bool(persistent)), int(blocks)),
              "schedule")

    def set_profiling(self, on):
        check(lib().ct_tracer_set_profiling(self.handle, int(bool(on))), "profiling")

    def set_counting(self, on):
        check(lib().ct_tracer_set_counting(self.handle, int(bool(on))), "counting")

    def accumulate_pass(self, ps, frame, width, height, frame_index, opts, d_accum,
                        stream=None):
        o = A.TraceOpts(int(opts.get("maxBounces", 8)), int(bool(opts.get("russianRoulette", True))),
                        int(bool(opts.get("nee", True))))
        fr = np.ascontiguousarray(frame, np.float64)
        check(lib().ct_accumulate_pass(self.handle, device_scene(ps).handle,
                                       fr.ctypes.data_as(C.c_void_p), width, height,
                                       frame_index & 0xFFFFFFFF, C.byref(o), d_accum.ptr, stream),
              "ct_accumulate_pass")

    def radiance(self, ps, d_rays, n, d_rng, opts, d_out, stream=None):
        """radianceInto (tracer.ts:191-328) of n rays on the device: d_rays
        (n*6 float64), d_rng (n uint32 RNG states, advanced in place), d_out
        (n*3 float64)."""
        o = A.TraceOpts(int(opts.get("maxBounces", 8)), int(bool(opts.get("russianRoulette", True))),
                        int(bool(opts.get("nee", True))))
        check(lib().ct_radiance(self.handle, device_scene(ps).handle, d_rays.ptr, int(n),
                                d_rng.ptr, C.byref(o), d_out.ptr, stream), "ct_radiance")

    def render_image(self, ps, frame, opts, d_out, stream=None):
        o = A.RenderOpts(opts["width"], opts["height"], opts["spp"], opts.get("seed", 0) & 0xFFFFFFFF,
                         int(opts.get("maxBounces", 8)), int(bool(opts.get("russianRoulette", True))),
                         int(bool(opts.get("nee", True))))
        fr = np.ascontiguousarray(frame, np.float64)
        check(lib().ct_render_image(self.handle, device_scene(ps).handle,
                                    fr.ctypes.data_as(C.c_void_p), C.byref(o), d_out.ptr, stream),
              "ct_render_image")

    def set_capacity(self, max_paths):
        check(lib().ct_tracer_set_capacity(self.handle, int(max_paths)), "set_capacity")

    def set_graph(self, on):
        check(lib().ct_tracer_set_graph(self.handle, int(bool(on))), "set_graph")

    def set_path_kernel(self, mode):
        """PATHS_WAVEFRONT (per-bounce launches), PATHS_PER_PATH (one launch, a
        lane per path at a time) or PATHS_AUTO (per-path for rounds of at most
        2^24 paths over scenes of at most 32768 primitives); the binary64
        results are identical under each."""
        check(lib().ct_tracer_set_path_kernel(self.handle, int(mode)), "set_path_kernel")

    def get_path_kernel(self):
        """(mode, rounds of the last call that ran the per-path launch)."""
        m, r = C.c_int32(0), C.c_uint32(0)
        check(lib().ct_tracer_get_path_kernel(self.handle, C.byref(m), C.byref(r)),
              "get_path_kernel")
        return m.value, r.value

    def set_overlap(self, on):
        """NEE launches beside the next extension launch (True) or in order."""
        check(lib().ct_tracer_set_overlap(self.handle, int(bool(on))), "set_overlap")

    def set_precision(self, precision):
        """PRECISION_F64 (reference arithmetic) or PRECISION_F32 (precision form)."""
        check(lib().ct_tracer_set_precision(self.handle, int(precision)), "set_precision")

    def set_workgroup(self, lanes):
        check(lib().ct_tracer_set_workgroup(self.handle, int(lanes)), "set_workgroup")

    def get_config(self):
        c = A.TracerConfig()
        check(lib().ct_tracer_get_config(self.handle, C.byref(c)), "ct_tracer_get_config")
        return {k: getattr(c, k) for k, _ in A.TracerConfig._fields_}

    def set_config(self, cfg):
        c = A.TracerConfig(**{k: int(v) for k, v in cfg.items()})
        check(lib().ct_tracer_set_config(self.handle, C.byref(c)), "ct_tracer_set_config")

    def autotune(self, jobs, max_trials=24, cache_path=None, stream=None):
        """Times render_batch(jobs) under candidate launch settings, applies the
        fastest and returns {config, ms, initialMs, trials, fromCache}; the jobs'
        outputs hold their images afterwards."""
        arr = self._jobs(jobs)
        r = A.TuneResult()
        path = None if cache_path is None else os.fsencode(cache_path)
        check(lib().ct_tracer_autotune(self.handle, C.cast(arr, C.c_void_p), len(jobs),
                                       int(max_trials), path, C.byref(r), stream),
              "ct_tracer_autotune")
        cfg = {k: getattr(r.config, k) for k, _ in A.TracerConfig._fields_}
        return {"config": cfg, "ms": r.ms, "initialMs": r.initialMs, "trials": r.trials,
                "fromCache": bool(r.fromCache)}

    @staticmethod
    def _frame16(frame):
        return (C.c_double * 16)(*[float(x) for x in frame])

    def render_batch(self, jobs, stream=None):
        """jobs: list of dicts {scene (PackedScene/DeviceScene), frame, width, height, spp,
        seed, maxBounces, russianRoulette, nee, out (DeviceBuffer)} -- one renderImage each,
        all traced in shared launches."""
        arr = self._jobs(jobs)
        check(lib().ct_render_batch(self.handle, C.cast(arr, C.c_void_p), len(jobs), stream),
              "ct_render_batch")

    def _jobs(self, jobs):
        arr = (A.Job * len(jobs))()
        for k, j in enumerate(jobs):
            arr[k].scene = device_scene(j["scene"]).handle.value
            arr[k].frame = self._frame16(j["frame"])
            arr[k].width, arr[k].height = j["width"], j["height"]
            arr[k].spp, arr[k].seed = j["spp"], j.get("seed", 0) & 0xFFFFFFFF
            arr[k].maxBounces = int(j.get("maxBounces", 8))
            arr[k].russianRoulette = int(bool(j.get("russianRoulette", True)))
            arr[k].nee = int(bool(j.get("nee", True)))
            arr[k].d_out = j["out"].ptr.value
        return arr

    def accumulate_batch(self, jobs, stream=None):
        """jobs: list of dicts {scene, frame, width, height, frameIndex, passes, maxBounces,
        russianRoulette, nee, accum (DeviceBuffer)} -- progressive passes, shared launches."""
        arr = (A.PassJob * len(jobs))()
        for k, j in enumerate(jobs):
            arr[k].scene = device_scene(j["scene"]).handle.value
            arr[k].frame = self._frame16(j["frame"])
            arr[k].width, arr[k].height = j["width"], j["height"]
            arr[k].frameIndex = j["frameIndex"] & 0xFFFFFFFF
            arr[k].passes = j["passes"]
            arr[k].maxBounces = int(j.get("maxBounces", 8))
            arr[k].russianRoulette = int(bool(j.get("russianRoulette", True)))
            arr[k].nee = int(bool(j.get("nee", True)))
            arr[k].d_accum = j["accum"].ptr.value
        check(lib().ct_accumulate_batch(self.handle, C.cast(arr, C.c_void_p), len(jobs), stream),
              "ct_accumulate_batch")

    def stats(self):
        s = A.RenderStats()
        check(lib().ct_tracer_stats(self.handle, C.byref(s)), "ct_tracer_stats")
        return {k: getattr(s, k) for k, _ in A.RenderStats._fields_}

    def close(self):
        if self.handle is not None and self.handle.value:
            lib().ct_tracer_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_tracer = None


def _tracer():
    global _default_tracer
    if _default_tracer is None:
        _default_tracer = Tracer()
    return _default_tracer


def accumulatePass(ps, frame, width, height, frameIndex, opts, accum):
    """accumulatePass (tracer.ts:361-395): adds one sample per pixel into the
    float32 sum `accum` (host numpy array, updated in place, or a DeviceBuffer)."""
    if isinstance(accum, DeviceBuffer):
        _tracer().accumulate_pass(ps, frame, width, height, frameIndex, opts, accum)
        return
    d = DeviceBuffer.from_array(np.ascontiguousarray(accum, np.float32))
    _tracer().accumulate_pass(ps, frame, width, height, frameIndex, opts, d)
    accum[...] = d.numpy().reshape(accum.shape)


def radiance_batch(ps, rays, rng_states, opts):
    """radianceInto (tracer.ts:191-328) of many rays: rays (n, 6) origin and
    direction, rng_states (n,) the xorshift32 states of their Rng objects.
    Returns (radiance (n, 3) float64, the states after the calls)."""
    rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
    n = len(rays)
    st = np.ascontiguousarray(rng_states, np.uint32).reshape(n)
    if n == 0:
        return np.zeros((0, 3)), st.copy()
    d_rays = DeviceBuffer.from_array(rays)
    d_rng = DeviceBuffer.from_array(st)
    d_out = DeviceBuffer.empty((n * 3,), np.float64)
    _tracer().radiance(ps, d_rays, n, d_rng, opts, d_out)
    return d_out.numpy().reshape(n, 3), d_rng.numpy()


def radianceInto(ps, ox, oy, oz, dx, dy, dz, rng, opts, out=None):
    """radianceInto (tracer.ts:191-328): the radiance along one ray, drawing
    from `rng` (an Rng; its state advances as the reference's does), into
    `out` (3 floats, a new array when omitted), which is returned."""
    L, st = radiance_batch(ps, [[ox, oy, oz, dx, dy, dz]], [rng.state], opts)
    rng.state = int(st[0])
    if out is None:
        out = np.zeros(3)
    out[0], out[1], out[2] = L[0]
    return out


def renderImage(ps, frame, opts):
    """renderImage (tracer.ts:348-355): linear RGB float32 (h*w*3), divided by spp."""
    n = opts["width"] * opts["height"] * 3
    d = DeviceBuffer.empty((n,), np.float32)
    _tracer().render_image(ps, frame, opts, d)
    return d.numpy()
