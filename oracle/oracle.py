"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, always as the checker, never as the product.

The scene format is the reference's plain-object scene (src/tracer/scene.ts:5-74)
written as Python dicts with tuples for Vec3/Color:

    {"objects": [{"name": str,
                  "shape": {"kind": "sphere", "center": (x,y,z), "radius": r}
                         | {"kind": "triangle", "a":..., "b":..., "c":...}
                         | {"kind": "quad", "origin":..., "u":..., "v":...}
                         | {"kind": "mesh", "vertices": (V,3) array, "indices": (3T,) array},
                  "material": {"type": "lambertian", "albedo": rgb,
                               "checker": {"color2": rgb, "scale": s}?, "emission": rgb?}
                            | {"type": "metal", "albedo": rgb, "roughness": r}
                            | {"type": "dielectric", "ior": n, "tint": rgb?}
                            | {"type": "emissive", "color": rgb, "strength": s}}],
     "camera": {"target", "yaw", "pitch", "distance", "fovDeg", "lensRadius", "focusDistance"},
     "sky": {"kind": "none"} | {"kind": "gradient", "zenith", "horizon", "strength"}}

The expansion below restates scene.ts:85-118 and pack.ts:75-192 independently of
the product's packer so that a packing mistake on either side shows up as a
parity failure.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

PRIM_STRIDE = 16
MAT_STRIDE = 16
EXACT_LEAF_LIMIT = 1 << 21  # float32 leaf words start*8+count are exact below this
NODE_STRIDE = 8
T_MIN = 1e-4


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: run `make -C {_HERE}`")
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


# Planted-mistake variants (o64.h, O64_PLANT_*): the same source built with
# -DO64_PLANT=k, used only by the tests that show each pin catches a mistake.
PLANTS = {"sah_le": 1, "axis_reversed": 2, "nudge_inward": 3, "partition_lt": 4,
          "median_unstable": 5, "median_axis_last": 6, "sweep_shift": 7, "leaf3": 8,
          "cone_no_sqrt": 20, "cone_angle_uniform": 21, "cone_pdf_2x": 22,
          "inside_half_area": 23, "lens_linear_r": 30, "lens_no_refocus": 31, "sky_flip": 40}
_CFLAGS = ["-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
           "-fexcess-precision=standard"]


def build_planted(names, outdir):
    """Builds liboracle_<name>.so for each planted mistake into outdir (in
    parallel); returns {name: path}."""
    import subprocess
    from concurrent.futures import ThreadPoolExecutor

    def one(name):
        out = os.path.join(outdir, f"liboracle_{name}.so")
        subprocess.run(["gcc", *_CFLAGS, f"-DO64_PLANT={PLANTS[name]}", "-shared", "-o", out,
                        os.path.join(_HERE, "o64.c"), "-lm"], check=True)
        return name, out

    with ThreadPoolExecutor(max_workers=8) as ex:
        return dict(ex.map(one, names))


class using_library:
    """Context in which every oracle function calls the library at `path`."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = lib()
        L = C.CDLL(self.path)
        _declare(L)
        _lib = L
        return L

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _p(t):
    return C.POINTER(t)


class O64Scene(C.Structure):
    _fields_ = [
        ("prims", _p(C.c_float)), ("primCount", C.c_int),
        ("nodes", _p(C.c_float)), ("nodeCount", C.c_int),
        ("mats", _p(C.c_float)), ("matCount", C.c_int),
        ("lights", _p(C.c_int32)), ("lightCount", C.c_int),
        ("primObject", _p(C.c_uint32)),
        ("skyKind", C.c_int),
        ("zenith", C.c_double * 3), ("horizon", C.c_double * 3), ("skyStrength", C.c_double),
        ("leafWords", _p(C.c_uint32)),
    ]


class O64TraceOpts(C.Structure):
    _fields_ = [("maxBounces", C.c_int), ("russianRoulette", C.c_int), ("nee", C.c_int)]


class O64Rng(C.Structure):
    _fields_ = [("state", C.c_uint32)]


def _declare(L):
    d, f, i32, u32, u8, i64 = C.c_double, C.c_float, C.c_int32, C.c_uint32, C.c_uint8, C.c_int64
    vp = C.c_void_p
    L.o64_pcg_hash.restype = u32
    L.o64_pcg_hash.argtypes = [u32]
    L.o64_seed_for.restype = u32
    L.o64_seed_for.argtypes = [u32, u32]
    L.o64_rng_stream.argtypes = [u32, C.c_int, vp]
    L.o64_rng_stream_u32.argtypes = [u32, C.c_int, vp]
    L.o64_pack_scene.restype = C.c_int
    L.o64_pack_scene.argtypes = [C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, _p(C.c_int)]
    L.o64_prim_bounds.argtypes = [C.c_int, vp, vp, vp, vp, vp]
    L.o64_pack_scene_ex.restype = C.c_int
    L.o64_pack_scene_ex.argtypes = [C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp,
                                    _p(C.c_int), vp]
    L.o64_build_bvh.restype = C.c_int
    L.o64_build_bvh.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp]
    L.o64_build_bvh_ex.restype = C.c_int
    L.o64_build_bvh_ex.argtypes = [C.c_int, vp, vp, vp, C.c_int, vp, vp, vp]
    L.o64_validate_bvh.restype = C.c_int
    L.o64_validate_bvh.argtypes = [vp, C.c_int, vp, C.c_int, vp]
    L.o64_validate_bvh_ex.restype = C.c_int
    L.o64_validate_bvh_ex.argtypes = [vp, C.c_int, vp, C.c_int, vp, vp]
    L.o64_partition_flags.restype = C.c_int
    L.o64_partition_flags.argtypes = [C.c_int, vp, vp]
    L.o64_closest_batch.argtypes = [_p(O64Scene), C.c_int, vp, d, vp, C.c_int, vp, vp, vp, vp, vp]
    L.o64_occluded_batch.argtypes = [_p(O64Scene), C.c_int, vp, d, vp, vp, vp]
    for name in ("o64_schlick", "o64_dielectric_reflectance"):
        getattr(L, name).restype = d
        getattr(L, name).argtypes = [d, d, d]
    L.o64_power_heuristic.restype = d
    L.o64_power_heuristic.argtypes = [d, d]
    L.o64_cosine_hemisphere.argtypes = [d, d, d, d, d, vp]
    L.o64_unit_vector.argtypes = [d, d, vp]
    L.o64_metal.argtypes = [d, d, d, d, d, d, d, d, d, d, vp]
    L.o64_dielectric.restype = C.c_int
    L.o64_dielectric.argtypes = [d, d, d, d, d, d, d, C.c_int, d, vp]
    L.o64_checker_primary.restype = C.c_int
    L.o64_checker_primary.argtypes = [d, d, d, d]
    L.o64_camera_frame.argtypes = [vp, d, d, d, d, d, d, d, vp]
    L.o64_generate_ray.argtypes = [vp, d, d, d, d, vp, vp]
    L.o64_accumulate_pass.argtypes = [_p(O64Scene), vp, C.c_int, C.c_int, u32, _p(O64TraceOpts),
                                      vp, C.c_int, C.c_int, vp]
    L.o64_pass_radiance.argtypes = [_p(O64Scene), vp, C.c_int, C.c_int, u32, _p(O64TraceOpts),
                                    vp, C.c_int, C.c_int]
    L.o64_render_image.argtypes = [_p(O64Scene), vp, C.c_int, C.c_int, C.c_int, u32,
                                   _p(O64TraceOpts), vp]
    for name in ("o64_aces", "o64_reinhard", "o64_linear_to_srgb"):
        getattr(L, name).restype = d
        getattr(L, name).argtypes = [d]
    L.o64_tonemap_buffer.argtypes = [vp, C.c_int, C.c_int, d, d, C.c_int, vp]
    for name in ("o64_js_sin", "o64_js_cos", "o64_js_cbrt"):
        getattr(L, name).restype = d
        getattr(L, name).argtypes = [d]
    L.o64_render_pixels.argtypes = [_p(O64Scene), vp, C.c_int, C.c_int, vp, C.c_int, C.c_int,
                                    u32, _p(O64TraceOpts), vp]
    L.o64_camera_sample.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, u32, vp, vp]
    L.o64_sample_light.restype = C.c_int
    L.o64_sample_light.argtypes = [_p(O64Scene), d, d, d, _p(u32), vp]
    L.o64_light_pdf.restype = d
    L.o64_light_pdf.argtypes = [_p(O64Scene), d, d, d, d, d, d, d, d, d, C.c_int]
    L.o64_radiance.argtypes = [_p(O64Scene), d, d, d, d, d, d, _p(O64Rng), _p(O64TraceOpts), vp,
                               vp]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# rng.ts

def pcg_hash(x):
    return lib().o64_pcg_hash(x & 0xFFFFFFFF)


def seed_for(pixel, frame):
    return lib().o64_seed_for(pixel & 0xFFFFFFFF, frame & 0xFFFFFFFF)


def rng_stream(seed, n):
    out = np.empty(n, np.float64)
    lib().o64_rng_stream(seed & 0xFFFFFFFF, n, _ptr(out))
    return out


def rng_stream_u32(seed, n):
    out = np.empty(n, np.uint32)
    lib().o64_rng_stream_u32(seed & 0xFFFFFFFF, n, _ptr(out))
    return out


class Rng:
    """Python view of rng.ts:25-65 driven by a pre-drawn stream from the oracle."""

    def __init__(self, seed, chunk=1 << 16):
        self._seed = seed
        self._buf = rng_stream(seed, chunk)
        self._pos = 0
        self._chunk = chunk

    def next(self):
        if self._pos >= len(self._buf):
            self._chunk *= 2
            self._buf = rng_stream(self._seed, self._chunk)
        v = self._buf[self._pos]
        self._pos += 1
        return float(v)

    def range(self, lo, hi):
        return lo + (hi - lo) * self.next()

    def take(self, n):
        """Next n draws as an array (same stream as n calls to next())."""
        need = self._pos + n
        if need > len(self._buf):
            self._chunk = max(self._chunk * 2, need)
            self._buf = rng_stream(self._seed, self._chunk)
        out = self._buf[self._pos:need].copy()
        self._pos = need
        return out


# ---------------------------------------------------------------------------
# scene.ts / pack.ts restated (expansion and material records)

def _vadd(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def expand_scene(scene):
    """Primitive list in scene order: kind (n,), geom (n,9) f64, obj (n,)."""
    kinds, geoms, objs = [], [], []
    for oi, ob in enumerate(scene["objects"]):
        sh = ob["shape"]
        k = sh["kind"]
        if k == "sphere":
            c = sh["center"]
            kinds.append(np.array([0], np.int32))
            geoms.append(np.array([[c[0], c[1], c[2], sh["radius"], 0, 0, 0, 0, 0]], np.float64))
            objs.append(np.array([oi], np.int32))
        elif k == "triangle":
            a, b, c = sh["a"], sh["b"], sh["c"]
            kinds.append(np.array([1], np.int32))
            geoms.append(np.array([[*a, *b, *c]], np.float64))
            objs.append(np.array([oi], np.int32))
        elif k == "quad":
            # scene.ts:85-94: p0 = o, p1 = o+u, p2 = (o+u)+v, p3 = o+v
            o, u, v = sh["origin"], sh["u"], sh["v"]
            p1 = _vadd(o, u)
            p2 = _vadd(_vadd(o, u), v)
            p3 = _vadd(o, v)
            kinds.append(np.array([1, 1], np.int32))
            geoms.append(np.array([[*o, *p1, *p2], [*o, *p2, *p3]], np.float64))
            objs.append(np.array([oi, oi], np.int32))
        elif k == "mesh":
            V = np.asarray(sh["vertices"], np.float64).reshape(-1, 3)
            I = np.asarray(sh["indices"], np.int64)
            t = len(I) // 3
            I = I[: 3 * t].reshape(t, 3)
            g = np.concatenate([V[I[:, 0]], V[I[:, 1]], V[I[:, 2]]], axis=1)
            kinds.append(np.ones(t, np.int32))
            geoms.append(g)
            objs.append(np.full(t, oi, np.int32))
        else:
            raise ValueError(k)
    if not kinds:
        return (np.zeros(0, np.int32), np.zeros((0, 9), np.float64), np.zeros(0, np.int32))
    return (np.ascontiguousarray(np.concatenate(kinds)),
            np.ascontiguousarray(np.concatenate(geoms)),
            np.ascontiguousarray(np.concatenate(objs)))


_MAT_CODE = {"lambertian": 0, "metal": 1, "dielectric": 2, "emissive": 3}


def material_emission(m):
    """scene.ts:176-185"""
    if m["type"] == "emissive":
        c, s = m["color"], m["strength"]
        return (c[0] * s, c[1] * s, c[2] * s)
    if m["type"] == "lambertian" and m.get("emission") is not None:
        e = m["emission"]
        if e[0] > 0 or e[1] > 0 or e[2] > 0:
            return e
    return None


def material_record(m):
    """pack.ts:75-119 (writeMaterial), one 16-float record."""
    o = np.zeros(MAT_STRIDE, np.float32)
    o[3] = _MAT_CODE[m["type"]]
    t = m["type"]
    if t == "lambertian":
        o[0:3] = m["albedo"]
        if m.get("checker"):
            o[5] = m["checker"]["scale"]
            o[6] = 1
            o[8:11] = m["checker"]["color2"]
    elif t == "metal":
        o[0:3] = m["albedo"]
        o[4] = m["roughness"]
    elif t == "dielectric":
        o[0:3] = m.get("tint") or (1, 1, 1)
        o[4] = m["ior"]
    elif t == "emissive":
        o[0:3] = m["color"]
        o[4] = m["strength"]
    e = material_emission(m)
    if e is not None:
        o[12:15] = e
    return o


class Packed:
    """The PackedScene of pack.ts:46-59 plus the oracle's view struct."""

    def __init__(self, prims, nodes, nodeCount, mats, lights, primObject, order, sky,
                 leafWords=None):
        self.prims = prims
        self.leafWords = leafWords
        self.primCount = len(order)
        self.nodes = nodes
        self.nodeCount = nodeCount
        self.mats = mats
        self.matCount = len(mats) // MAT_STRIDE
        self.lights = lights
        self.primObject = primObject
        self.order = order
        self.sky = sky
        s = O64Scene()
        s.prims = prims.ctypes.data_as(_p(C.c_float))
        s.primCount = self.primCount
        s.nodes = nodes.ctypes.data_as(_p(C.c_float))
        s.nodeCount = nodeCount
        s.mats = mats.ctypes.data_as(_p(C.c_float))
        s.matCount = self.matCount
        s.lights = lights.ctypes.data_as(_p(C.c_int32))
        s.lightCount = len(lights)
        s.primObject = primObject.ctypes.data_as(_p(C.c_uint32))
        if sky.get("kind") == "gradient":
            s.skyKind = 1
            s.zenith = (C.c_double * 3)(*sky["zenith"])
            s.horizon = (C.c_double * 3)(*sky["horizon"])
            s.skyStrength = sky["strength"]
        else:
            s.skyKind = 0
        if leafWords is not None:  # the exact-leaf-word layout extension
            s.leafWords = leafWords.ctypes.data_as(_p(C.c_uint32))
        self.view = s


def pack_scene(scene, bins=12):
    """packScene (pack.ts:195-223) computed by the oracle."""
    kind, geom, obj = expand_scene(scene)
    n = len(kind)
    cap = max(1, n)
    objs = scene["objects"]
    mats = np.zeros(max(1, len(objs)) * MAT_STRIDE, np.float32)
    emits = np.zeros(max(1, len(objs)), np.uint8)
    for i, ob in enumerate(objs):
        mats[i * MAT_STRIDE:(i + 1) * MAT_STRIDE] = material_record(ob["material"])
        emits[i] = material_emission(ob["material"]) is not None
    prims = np.zeros(cap * PRIM_STRIDE, np.float32)
    primObject = np.zeros(cap, np.uint32)
    nodes = np.zeros(max(1, 2 * n - 1) * NODE_STRIDE, np.float32)
    order = np.zeros(cap, np.uint32)
    lights = np.zeros(cap, np.int32)
    nl = C.c_int(0)
    # from 2^21 primitives the float32 leaf words (bvh.ts:60-62) are inexact:
    # the exact words travel beside the nodes, as in the product's packScene
    words = np.zeros(max(1, 2 * n - 1), np.uint32) if n >= EXACT_LEAF_LIMIT else None
    nc = lib().o64_pack_scene_ex(n, _ptr(kind), _ptr(geom), _ptr(obj), _ptr(emits), bins,
                                 _ptr(prims), _ptr(primObject), _ptr(nodes), _ptr(order),
                                 _ptr(lights), C.byref(nl), _ptr(words) if words is not None
                                 else None)
    return Packed(prims, nodes[: nc * NODE_STRIDE].copy(), nc, mats,
                  lights[: nl.value].copy(), primObject, order[:n].copy(),
                  scene.get("sky", {"kind": "none"}),
                  words[:nc].copy() if words is not None else None)


def prim_bounds(kind, geom):
    n = len(kind)
    bmin = np.zeros(max(1, n) * 3)
    bmax = np.zeros(max(1, n) * 3)
    bcen = np.zeros(max(1, n) * 3)
    lib().o64_prim_bounds(n, _ptr(np.ascontiguousarray(kind, np.int32)),
                          _ptr(np.ascontiguousarray(geom, np.float64)), _ptr(bmin), _ptr(bmax),
                          _ptr(bcen))
    return bmin[: 3 * n], bmax[: 3 * n], bcen[: 3 * n]


def build_bvh(bmin, bmax, bcen, bins=12, leaf_words=False):
    """buildBvh (bvh.ts:299) on PrimBounds arrays. Returns (nodes, nodeCount,
    order), plus the exact leaf word of every node when leaf_words is set."""
    n = len(bmin) // 3
    nodes = np.zeros(max(1, 2 * n - 1) * NODE_STRIDE, np.float32)
    order = np.zeros(max(1, n), np.uint32)
    words = np.zeros(max(1, 2 * n - 1), np.uint32) if leaf_words else None
    nc = lib().o64_build_bvh_ex(n, _ptr(np.ascontiguousarray(bmin, np.float64)),
                                _ptr(np.ascontiguousarray(bmax, np.float64)),
                                _ptr(np.ascontiguousarray(bcen, np.float64)), bins, _ptr(nodes),
                                _ptr(order), _ptr(words) if leaf_words else None)
    if leaf_words:
        return nodes[: nc * NODE_STRIDE].copy(), nc, order[:n].copy(), words[:nc].copy()
    return nodes[: nc * NODE_STRIDE].copy(), nc, order[:n].copy()


def validate_bvh(nodes, nodeCount, order, primCount, leafWords=None):
    stats = np.zeros(3, np.int32)
    words = None if leafWords is None else np.ascontiguousarray(leafWords, np.uint32)
    errs = lib().o64_validate_bvh_ex(_ptr(np.ascontiguousarray(nodes, np.float32)), nodeCount,
                                     _ptr(np.ascontiguousarray(order, np.uint32)), primCount,
                                     _ptr(words) if words is not None else None, _ptr(stats))
    return {"errors": errs, "maxLeafSize": int(stats[0]), "leafCount": int(stats[1]),
            "depth": int(stats[2])}


def partition_flags(good):
    good = np.ascontiguousarray(good, np.uint8)
    perm = np.zeros(max(1, len(good)), np.int32)
    lo = lib().o64_partition_flags(len(good), _ptr(good), _ptr(perm))
    return perm[: len(good)], lo


# ---------------------------------------------------------------------------
# intersect.ts

WALK_ORDERED, WALK_STACKLESS, WALK_BRUTE = 0, 1, 2


def closest_batch(ps, rays, tmax=None, walk=WALK_ORDERED, tmin=T_MIN):
    rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
    n = len(rays)
    tm = np.full(n, np.inf) if tmax is None else np.ascontiguousarray(np.broadcast_to(tmax, (n,)), np.float64)
    prim = np.zeros(n, np.int32)
    t = np.zeros(n)
    u = np.zeros(n)
    v = np.zeros(n)
    counts = np.zeros(4, np.int64)
    lib().o64_closest_batch(C.byref(ps.view), n, _ptr(rays), tmin, _ptr(tm), walk, _ptr(prim),
                            _ptr(t), _ptr(u), _ptr(v), _ptr(counts))
    return {"prim": prim, "t": t, "u": u, "v": v, "counts": counts}


def occluded_batch(ps, rays, tmax, tmin=T_MIN):
    rays = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
    n = len(rays)
    tm = np.ascontiguousarray(np.broadcast_to(tmax, (n,)), np.float64)
    out = np.zeros(max(1, n), np.uint8)
    counts = np.zeros(4, np.int64)
    lib().o64_occluded_batch(C.byref(ps.view), n, _ptr(rays), tmin, _ptr(tm), _ptr(out),
                             _ptr(counts))
    return out[:n].astype(bool), counts


# ---------------------------------------------------------------------------
# materials.ts

def schlick(c, a, b):
    return lib().o64_schlick(c, a, b)


def dielectric_reflectance(c, a, b):
    return lib().o64_dielectric_reflectance(c, a, b)


def power_heuristic(a, b):
    return lib().o64_power_heuristic(a, b)


def cosine_hemisphere(n, u1, u2):
    out = np.zeros(3)
    lib().o64_cosine_hemisphere(n[0], n[1], n[2], u1, u2, _ptr(out))
    return out


def unit_vector(u1, u2):
    out = np.zeros(3)
    lib().o64_unit_vector(u1, u2, _ptr(out))
    return out


def metal(d, n, rough, u1, u2, u3):
    out = np.zeros(3)
    lib().o64_metal(d[0], d[1], d[2], n[0], n[1], n[2], rough, u1, u2, u3, _ptr(out))
    return out


def dielectric(d, n, ior, front, u):
    out = np.zeros(3)
    refl = lib().o64_dielectric(d[0], d[1], d[2], n[0], n[1], n[2], ior, int(front), u, _ptr(out))
    return out, bool(refl)


def checker_primary(p, cells):
    return bool(lib().o64_checker_primary(p[0], p[1], p[2], cells))


# ---------------------------------------------------------------------------
# camera.ts

def camera_frame(cam, aspect):
    f = np.zeros(16)
    tgt = np.array(cam["target"], np.float64)
    lib().o64_camera_frame(_ptr(tgt), cam["yaw"], cam["pitch"], cam["distance"], cam["fovDeg"],
                           cam["lensRadius"], cam["focusDistance"], aspect, _ptr(f))
    return f


def generate_ray(frame, sx, sy, lx=0.0, ly=0.0):
    o = np.zeros(3)
    d = np.zeros(3)
    lib().o64_generate_ray(_ptr(np.ascontiguousarray(frame, np.float64)), sx, sy, lx, ly, _ptr(o),
                           _ptr(d))
    return o, d


def js_sin(x):
    """Math.sin as the JS engine computes it (fdlibm 5.3, jsmath.h)."""
    return lib().o64_js_sin(x)


def js_cos(x):
    return lib().o64_js_cos(x)


def js_cbrt(x):
    """Math.cbrt as the JS engine computes it (FreeBSD msun s_cbrt, jsmath.h)."""
    return lib().o64_js_cbrt(x)


def camera_sample(frame, width, height, x, y, frame_index):
    """The camera ray of the path at pixel (x, y) in pass frame_index
    (tracer.ts:373-383: jitter, lens draws, generateRay)."""
    o = np.zeros(3)
    d = np.zeros(3)
    lib().o64_camera_sample(_ptr(np.ascontiguousarray(frame, np.float64)), width, height, x, y,
                            frame_index & 0xFFFFFFFF, _ptr(o), _ptr(d))
    return o, d


# ---------------------------------------------------------------------------
# tracer.ts

def sample_light(ps, p, state):
    """sampleLightInto (tracer.ts:93-180) at point p from RNG state `state`.
    Returns (ok, [dir3, dist, pdf, emission3], state after the draws)."""
    st = C.c_uint32(state & 0xFFFFFFFF)
    out = np.zeros(8)
    ok = lib().o64_sample_light(C.byref(ps.view), p[0], p[1], p[2], C.byref(st), _ptr(out))
    return bool(ok), out, st.value


def light_pdf(ps, p, h, hn, prim):
    """lightPdfScalar (tracer.ts:61-90)."""
    return lib().o64_light_pdf(C.byref(ps.view), p[0], p[1], p[2], h[0], h[1], h[2], hn[0],
                               hn[1], hn[2], int(prim))


def radiance(ps, o, d, state, max_bounces=8, russian_roulette=True, nee=True,
             with_state=False):
    """radianceInto (tracer.ts:191-328) of one ray with RNG state `state`; with
    with_state also the RNG state after the call."""
    rng = O64Rng(state & 0xFFFFFFFF)
    out = np.zeros(3)
    opts = _opts(max_bounces, russian_roulette, nee)
    lib().o64_radiance(C.byref(ps.view), o[0], o[1], o[2], d[0], d[1], d[2], C.byref(rng),
                       C.byref(opts), _ptr(out), None)
    return (out, rng.state) if with_state else out


def _opts(max_bounces=8, russian_roulette=True, nee=True):
    return O64TraceOpts(int(max_bounces), int(bool(russian_roulette)), int(bool(nee)))


def accumulate_pass(ps, frame, width, height, frame_index, accum, max_bounces=8,
                    russian_roulette=True, nee=True, threads=1, rows=None):
    """accumulatePass (tracer.ts:361-395) into float32 accum (w*h*3), rows split over threads.

    Returns [extension rays, shadow rays] traced.
    """
    assert accum.dtype == np.float32 and accum.size == width * height * 3
    frame = np.ascontiguousarray(frame, np.float64)
    opts = _opts(max_bounces, russian_roulette, nee)
    y0, y1 = rows if rows is not None else (0, height)
    counts = [np.zeros(2, np.int64) for _ in range(max(1, threads))]

    def work(t):
        # rows interleaved in bands of one row per thread
        for y in range(y0 + t, y1, threads):
            lib().o64_accumulate_pass(C.byref(ps.view), _ptr(frame), width, height,
                                      frame_index & 0xFFFFFFFF, C.byref(opts), _ptr(accum), y,
                                      y + 1, _ptr(counts[t]))

    if threads <= 1:
        work(0) if rows is not None else lib().o64_accumulate_pass(
            C.byref(ps.view), _ptr(frame), width, height, frame_index & 0xFFFFFFFF,
            C.byref(opts), _ptr(accum), y0, y1, _ptr(counts[0]))
    else:
        ts = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
        for th in ts:
            th.start()
        for th in ts:
            th.join()
    return sum(counts)


def pass_radiance(ps, frame, width, height, frame_index, max_bounces=8, russian_roulette=True,
                  nee=True):
    """Per-path radiance (w*h, 3) of one pass, before accumulation."""
    out = np.zeros(width * height * 3)
    opts = _opts(max_bounces, russian_roulette, nee)
    lib().o64_pass_radiance(C.byref(ps.view), _ptr(np.ascontiguousarray(frame, np.float64)),
                            width, height, frame_index & 0xFFFFFFFF, C.byref(opts), _ptr(out),
                            0, height)
    return out.reshape(-1, 3)


def render_image(ps, frame, width, height, spp, seed=0, max_bounces=8, russian_roulette=True,
                 nee=True):
    """renderImage (tracer.ts:348-355): averaged linear RGB float32 (h*w*3)."""
    out = np.zeros(width * height * 3, np.float32)
    opts = _opts(max_bounces, russian_roulette, nee)
    lib().o64_render_image(C.byref(ps.view), _ptr(np.ascontiguousarray(frame, np.float64)),
                           width, height, spp, seed & 0xFFFFFFFF, C.byref(opts), _ptr(out))
    return out


def render_pixels(ps, frame, width, height, pixels, spp, seed=0, max_bounces=8,
                  russian_roulette=True, nee=True, threads=1):
    """renderImage's values (count, 3) at pixel indices y*width + x, the list
    split over threads."""
    pixels = np.ascontiguousarray(pixels, np.uint32)
    out = np.zeros((len(pixels), 3), np.float32)
    opts = _opts(max_bounces, russian_roulette, nee)
    fr = np.ascontiguousarray(frame, np.float64)
    chunks = np.array_split(np.arange(len(pixels)), max(1, threads))

    def work(idx):
        if len(idx) == 0:
            return
        sub = np.ascontiguousarray(pixels[idx])
        res = np.zeros((len(idx), 3), np.float32)
        lib().o64_render_pixels(C.byref(ps.view), _ptr(fr), width, height, _ptr(sub), len(idx),
                                spp, seed & 0xFFFFFFFF, C.byref(opts), _ptr(res))
        out[idx] = res

    ts = [threading.Thread(target=work, args=(c,)) for c in chunks]
    for th in ts:
        th.start()
    for th in ts:
        th.join()
    return out


# ---------------------------------------------------------------------------
# tonemap.ts

JS_BENCH = os.path.join(_HERE, "js", "cpu_bench.js")


def node_binary():
    """Path of the `node` executable, or None."""
    import shutil
    return shutil.which("node")


def render_band_js(ps, frame, width, height, spp, seed, rows, threads, trials=1,
                   max_bounces=8, russian_roulette=True, nee=True, workdir=None, timeout=None):
    """renderImage's passes over rows [y0, y1) by the JS restatement of the
    reference tracer (oracle/js/trace.js) on `threads` Node worker threads,
    `trials` timed runs after a warm-up.  Returns (report, sums): the report
    of cpu_bench.js (seconds and rays per trial, thread count, CPU model) and
    the float32 per-pixel sums of the band's passes (w*h*3, zero outside the
    band) -- not yet times 1/spp."""
    import json
    import subprocess
    import tempfile

    node = node_binary()
    if node is None:
        raise RuntimeError("render_band_js: no node executable")
    own = workdir is None
    d = tempfile.mkdtemp(prefix="ct_js_") if own else workdir
    try:
        arrays = {"prims": ("f32", ps.prims), "nodes": ("f32", ps.nodes),
                  "mats": ("f32", ps.mats), "lights": ("i32", ps.lights),
                  "primObject": ("u32", ps.primObject)}
        if ps.leafWords is not None:
            arrays["leafWords"] = ("u32", ps.leafWords)
        dtypes = {"f32": np.float32, "i32": np.int32, "u32": np.uint32}
        for name, (t, a) in arrays.items():
            np.ascontiguousarray(a, dtypes[t]).tofile(os.path.join(d, f"{name}.{t}"))
        sky = ps.sky if ps.sky.get("kind") == "gradient" else {"kind": "none"}
        if sky.get("kind") == "gradient":
            sky = {"kind": "gradient", "zenith": [float(x) for x in sky["zenith"]],
                   "horizon": [float(x) for x in sky["horizon"]],
                   "strength": float(sky["strength"])}
        job = {"arrays": {k: v[0] for k, v in arrays.items()}, "nodeCount": int(ps.nodeCount),
               "lightCount": int(len(ps.lights)), "sky": sky,
               "frame": [float(x) for x in frame], "width": int(width), "height": int(height),
               "spp": int(spp), "seed": int(seed) & 0xFFFFFFFF, "maxBounces": int(max_bounces),
               "russianRoulette": bool(russian_roulette), "nee": bool(nee),
               "rows": [int(rows[0]), int(rows[1])], "threads": int(threads),
               "trials": int(trials)}
        with open(os.path.join(d, "job.json"), "w") as f:
            json.dump(job, f)
        r = subprocess.run([node, JS_BENCH, d], capture_output=True, text=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"cpu_bench.js failed ({r.returncode}): {r.stderr[-2000:]}")
        report = json.loads(r.stdout.strip().splitlines()[-1])
        sums = np.fromfile(os.path.join(d, "accum.f32"), np.float32)
        return report, sums
    finally:
        if own:
            import shutil
            shutil.rmtree(d, ignore_errors=True)


def aces(x):
    return lib().o64_aces(x)


def reinhard(x):
    return lib().o64_reinhard(x)


def linear_to_srgb(x):
    return lib().o64_linear_to_srgb(x)


def tonemap_buffer(accum, width, height, samples, exposure_ev=0.0, mode="aces"):
    rgba = np.zeros(width * height * 4, np.uint8)
    lib().o64_tonemap_buffer(_ptr(np.ascontiguousarray(accum, np.float32)), width, height,
                             float(samples), float(exposure_ev), 0 if mode == "aces" else 1,
                             _ptr(rgba))
    return rgba


def camera_at(eye, target, fov_deg, lens_radius, focus_distance=None):
    """cameraAt (scenes.ts:26-41)."""
    dx, dy, dz = eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]
    distance = math.sqrt(dx * dx + dy * dy + dz * dz)
    return {"target": tuple(target), "yaw": math.atan2(dx, dz), "pitch": math.asin(dy / distance),
            "distance": distance, "fovDeg": fov_deg, "lensRadius": lens_radius,
            "focusDistance": distance if focus_distance is None else focus_distance}
