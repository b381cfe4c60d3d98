"""ctypes binding of libcaustic_amd.so (the C ABI declared in include/caustic_amd.h)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcaustic_amd.so")

_lib = None


class CTError(RuntimeError):
    pass


class Sky(C.Structure):
    _fields_ = [("kind", C.c_int32), ("zenith", C.c_double * 3), ("horizon", C.c_double * 3),
                ("strength", C.c_double)]


class TraceOpts(C.Structure):
    _fields_ = [("maxBounces", C.c_int32), ("russianRoulette", C.c_int32), ("nee", C.c_int32)]


class RenderOpts(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("seed", C.c_uint32), ("maxBounces", C.c_int32), ("russianRoulette", C.c_int32),
                ("nee", C.c_int32)]


class RenderStats(C.Structure):
    _fields_ = [("extensionRays", C.c_uint64), ("shadowRays", C.c_uint64), ("paths", C.c_uint64),
                ("extendLaunches", C.c_uint32), ("extendMs", C.c_double), ("shadeMs", C.c_double),
                ("shadowMs", C.c_double), ("nodeVisits", C.c_uint64), ("triTests", C.c_uint64),
                ("sphereTests", C.c_uint64), ("shadowNodeVisits", C.c_uint64),
                ("shadowTriTests", C.c_uint64), ("shadowSphereTests", C.c_uint64),
                ("stackFallbacks", C.c_uint64), ("shadowStackFallbacks", C.c_uint64),
                ("rounds", C.c_uint32)]


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("frame", C.c_double * 16), ("width", C.c_uint32),
                ("height", C.c_uint32), ("spp", C.c_uint32), ("seed", C.c_uint32),
                ("maxBounces", C.c_int32), ("russianRoulette", C.c_int32), ("nee", C.c_int32),
                ("d_out", C.c_void_p)]


class PassJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("frame", C.c_double * 16), ("width", C.c_uint32),
                ("height", C.c_uint32), ("frameIndex", C.c_uint32), ("passes", C.c_uint32),
                ("maxBounces", C.c_int32), ("russianRoulette", C.c_int32), ("nee", C.c_int32),
                ("d_accum", C.c_void_p)]


class TracerConfig(C.Structure):
    _fields_ = [("maxPaths", C.c_uint64), ("walk", C.c_int32), ("shadowWalk", C.c_int32),
                ("stackDepth", C.c_int32), ("persistent", C.c_int32), ("blocks", C.c_int32),
                ("workgroup", C.c_int32), ("pathGroups", C.c_int32)]


class TuneResult(C.Structure):
    _fields_ = [("config", TracerConfig), ("ms", C.c_double), ("initialMs", C.c_double),
                ("trials", C.c_uint32), ("fromCache", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("hbmBudget", C.c_uint64), ("maxPaths", C.c_uint64), ("maxActive", C.c_uint32),
                ("slicePasses", C.c_uint32), ("policy", C.c_int32), ("graph", C.c_int32),
                ("precision", C.c_int32)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("passesDone", C.c_uint32), ("spp", C.c_uint32),
                ("submitSec", C.c_double), ("admitSec", C.c_double), ("doneSec", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("queued", C.c_uint64), ("active", C.c_uint64), ("rounds", C.c_uint64),
                ("paths", C.c_uint64), ("extensionRays", C.c_uint64), ("shadowRays", C.c_uint64),
                ("pixelSamples", C.c_uint64), ("hbmInUse", C.c_uint64), ("busySec", C.c_double),
                ("nowSec", C.c_double)]


# Every function of the C ABI with its argument types (restype is int status
# unless listed in _RESTYPES).
vp, u32, i32, f64, sz = C.c_void_p, C.c_uint32, C.c_int32, C.c_double, C.c_size_t
SIGNATURES = {
    "ct_init": [i32],
    "ct_last_error": [],
    "ct_device_alloc": [C.POINTER(vp), sz],
    "ct_device_free": [vp],
    "ct_copy_h2d": [vp, vp, sz],
    "ct_copy_d2h": [vp, vp, sz],
    "ct_stream_sync": [vp],
    "ct_scene_create": [vp, u32, vp, u32, vp, u32, vp, u32, vp, C.POINTER(Sky), C.POINTER(vp)],
    "ct_scene_create_device": [vp, u32, vp, u32, vp, u32, vp, u32, vp, C.POINTER(Sky), vp,
                               C.POINTER(vp)],
    "ct_scene_create_ex": [vp, u32, vp, u32, vp, u32, vp, u32, vp, C.POINTER(Sky), vp,
                           C.POINTER(vp)],
    "ct_scene_destroy": [vp],
    "ct_closest_hit": [vp, vp, u32, f64, vp, i32, vp, vp, vp, vp, vp],
    "ct_occluded": [vp, vp, u32, f64, vp, vp, vp],
    "ct_closest_hit_counts": [vp, vp, u32, f64, vp, i32, C.POINTER(C.c_uint64), vp],
    "ct_tracer_create": [C.POINTER(vp)],
    "ct_tracer_destroy": [vp],
    "ct_tracer_set_profiling": [vp, i32],
    "ct_tracer_set_counting": [vp, i32],
    "ct_tracer_set_walk": [vp, i32],
    "ct_tracer_set_traversal": [vp, i32, i32],
    "ct_tracer_set_schedule": [vp, i32, i32],
    "ct_accumulate_pass": [vp, vp, vp, u32, u32, u32, C.POINTER(TraceOpts), vp, vp],
    "ct_render_image": [vp, vp, vp, C.POINTER(RenderOpts), vp, vp],
    "ct_radiance": [vp, vp, vp, u32, vp, C.POINTER(TraceOpts), vp, vp],
    "ct_tracer_stats": [vp, C.POINTER(RenderStats)],
    "ct_render_batch": [vp, vp, u32, vp],
    "ct_accumulate_batch": [vp, vp, u32, vp],
    "ct_tracer_set_capacity": [vp, C.c_uint64],
    "ct_tracer_set_graph": [vp, i32],
    "ct_tracer_set_overlap": [vp, i32],
    "ct_tracer_set_path_kernel": [vp, i32],
    "ct_tracer_get_path_kernel": [vp, C.POINTER(i32), C.POINTER(u32)],
    "ct_bvh_build": [vp, vp, vp, u32, i32, vp, C.POINTER(u32), vp, vp],
    "ct_bvh_build_ex": [vp, vp, vp, u32, i32, vp, C.POINTER(u32), vp, vp, vp],
    "ct_bvh_build_noalloc": [vp, vp, vp, u32, i32, vp, C.POINTER(u32), vp, vp, vp],
    "ct_tracer_set_workgroup": [vp, i32],
    "ct_tracer_set_precision": [vp, i32],
    "ct_closest_hit32": [vp, vp, u32, C.c_float, vp, i32, vp, vp, vp, vp, vp],
    "ct_occluded32": [vp, vp, u32, C.c_float, vp, i32, vp, vp],
    "ct_closest_hit32_counts": [vp, vp, u32, C.c_float, vp, vp, vp],
    "ct_image_sum_f64": [vp, C.c_uint64, vp, vp],
    "ct_tonemap": [vp, u32, u32, f64, f64, i32, vp, vp],
    "ct_job_read_rgba": [vp, C.c_uint64, f64, i32, vp, i32],
    "ct_tracer_get_config": [vp, C.POINTER(TracerConfig)],
    "ct_tracer_set_config": [vp, C.POINTER(TracerConfig)],
    "ct_tracer_autotune": [vp, vp, u32, u32, C.c_char_p, C.POINTER(TuneResult), vp],
    "ct_server_create": [C.POINTER(ServerOpts), C.POINTER(vp)],
    "ct_server_destroy": [vp],
    "ct_job_submit": [vp, vp, C.POINTER(C.c_uint64)],
    "ct_job_poll": [vp, C.c_uint64, C.POINTER(JobStatus)],
    "ct_job_wait": [vp, C.c_uint64, f64, C.POINTER(JobStatus)],
    "ct_job_read": [vp, C.c_uint64, vp, i32],
    "ct_job_release": [vp, C.c_uint64],
    "ct_server_stats_get": [vp, C.POINTER(ServerStats)],
}
_RESTYPES = {"ct_last_error": C.c_char_p}


def lib():
    """Load the HIP library; fails loudly when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CTError(f"{LIB_PATH} is missing: build it with caustic-tracer_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, C.c_int)
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().ct_last_error().decode(errors="replace")
        raise CTError(f"{what} failed ({rc}): {msg}")


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        check(lib().ct_init(device), "ct_init")
        _initialised = True


class DeviceBuffer:
    """A device (HBM) allocation holding one numpy-typed array."""

    def __init__(self, nbytes, dtype=np.uint8, shape=None):
        init()
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.shape = shape if shape is not None else (self.nbytes // self.dtype.itemsize,)
        p = vp()
        check(lib().ct_device_alloc(C.byref(p), max(1, self.nbytes)), "ct_device_alloc")
        self.ptr = p

    @classmethod
    def empty(cls, shape, dtype):
        shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return cls(n, dtype, shape)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, a.dtype, a.shape)
        if a.nbytes:
            check(lib().ct_copy_h2d(b.ptr, a.ctypes.data_as(vp), a.nbytes), "ct_copy_h2d")
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes <= self.nbytes
        if a.nbytes:
            check(lib().ct_copy_h2d(self.ptr, a.ctypes.data_as(vp), a.nbytes), "ct_copy_h2d")

    def numpy(self, count=None):
        """The buffer on the host; with `count`, only its first `count` elements
        (a 1-D array)."""
        if count is None:
            out = np.empty(self.shape, self.dtype)
        else:
            total = int(np.prod(self.shape))
            if not 0 <= count <= total:
                raise ValueError(f"numpy(count={count}): the buffer holds {total} elements")
            out = np.empty((int(count),), self.dtype)
        if out.nbytes:
            check(lib().ct_copy_d2h(out.ctypes.data_as(vp), self.ptr, out.nbytes), "ct_copy_d2h")
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().ct_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync(stream=None):
    check(lib().ct_stream_sync(stream), "ct_stream_sync")
