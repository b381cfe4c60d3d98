"""Host packer: scene -> the packed typed arrays of pack.ts, BVH built on the GPU.

Scenes use the reference's plain-object format (scene.ts:5-74) as Python dicts
(see packScene).  The primitive records, normals, areas and PrimBounds are the
binary64 computations of pack.ts:131-192 / scene.ts:85-127, vectorised with
numpy (one correctly rounded operation per JS operation, same order); the BVH is
built by the GPU builder (ct_bvh_build) and prims are then reordered into leaf
order (pack.ts:195-204).
"""
import ctypes as C

import numpy as np

from . import _capi as A
from ._capi import DeviceBuffer, check, lib

PRIM_STRIDE = 16
MAT_STRIDE = 16
NODE_STRIDE = 8
_MAT_CODE = {"lambertian": 0, "metal": 1, "dielectric": 2, "emissive": 3}


class PackedScene:
    """PackedScene (pack.ts:46-59)."""

    def __init__(self, prims, primCount, nodes, nodeCount, mats, matCount, lights, primObject,
                 sky, order):
        self.prims = prims
        self.primCount = primCount
        self.nodes = nodes
        self.nodeCount = nodeCount
        self.mats = mats
        self.matCount = matCount
        self.lights = lights
        self.primObject = primObject
        self.sky = sky
        self.order = order  # bvh.order
        self.leafWords = None  # exact uint32 leaf words (layout extension) or None


def _expand(scene):
    """Primitives in scene order as (kind, a, b, c, radius, obj) arrays."""
    kinds, A3, B3, C3, R, objs = [], [], [], [], [], []
    for oi, ob in enumerate(scene["objects"]):
        sh = ob["shape"]
        k = sh["kind"]
        if k == "sphere":
            kinds.append(np.zeros(1, np.int32))
            A3.append(np.array([sh["center"]], np.float64))
            B3.append(np.zeros((1, 3)))
            C3.append(np.zeros((1, 3)))
            R.append(np.array([sh["radius"]], np.float64))
            objs.append(np.array([oi], np.int64))
        elif k == "triangle":
            kinds.append(np.ones(1, np.int32))
            A3.append(np.array([sh["a"]], np.float64))
            B3.append(np.array([sh["b"]], np.float64))
            C3.append(np.array([sh["c"]], np.float64))
            R.append(np.zeros(1))
            objs.append(np.array([oi], np.int64))
        elif k == "quad":
            o = np.array(sh["origin"], np.float64)
            u = np.array(sh["u"], np.float64)
            v = np.array(sh["v"], np.float64)
            p1 = o + u          # scene.ts:85-94
            p2 = (o + u) + v
            p3 = o + v
            kinds.append(np.ones(2, np.int32))
            A3.append(np.stack([o, o]))
            B3.append(np.stack([p1, p2]))
            C3.append(np.stack([p2, p3]))
            R.append(np.zeros(2))
            objs.append(np.array([oi, oi], np.int64))
        elif k == "mesh":
            V = np.asarray(sh["vertices"], np.float64).reshape(-1, 3)
            I = np.asarray(sh["indices"], np.int64)
            t = len(I) // 3
            I = I[: 3 * t].reshape(t, 3)
            kinds.append(np.ones(t, np.int32))
            A3.append(V[I[:, 0]])
            B3.append(V[I[:, 1]])
            C3.append(V[I[:, 2]])
            R.append(np.zeros(t))
            objs.append(np.full(t, oi, np.int64))
        else:
            raise ValueError(f"unknown shape kind {k!r}")
    if not kinds:
        z = np.zeros((0, 3))
        return np.zeros(0, np.int32), z, z, z, np.zeros(0), np.zeros(0, np.int64)
    return (np.concatenate(kinds), np.concatenate(A3), np.concatenate(B3), np.concatenate(C3),
            np.concatenate(R), np.concatenate(objs))


def material_emission(m):
    """materialEmission (scene.ts:176-185)."""
    if m["type"] == "emissive":
        c, s = m["color"], m["strength"]
        return (c[0] * s, c[1] * s, c[2] * s)
    if m["type"] == "lambertian" and m.get("emission") is not None:
        e = m["emission"]
        if e[0] > 0 or e[1] > 0 or e[2] > 0:
            return e
    return None


def write_material(out, i, m):
    """writeMaterial (pack.ts:75-119)."""
    o = out[i * MAT_STRIDE:(i + 1) * MAT_STRIDE]
    o[:] = 0
    o[3] = _MAT_CODE[m["type"]]
    t = m["type"]
    if t in ("lambertian", "metal"):
        o[0:3] = m["albedo"]
        if t == "metal":
            o[4] = m["roughness"]
        elif m.get("checker"):
            o[5] = m["checker"]["scale"]
            o[6] = 1
            o[8:11] = m["checker"]["color2"]
    elif t == "dielectric":
        o[0:3] = m.get("tint") or (1, 1, 1)
        o[4] = m["ior"]
    else:
        o[0:3] = m["color"]
        o[4] = m["strength"]
    e = material_emission(m)
    if e is not None:
        o[12:15] = e


def primBounds(kind, a, b, c, r):
    """PrimBounds as enumeratePrimitives computes them (pack.ts:150-187):
    (min, max, centroid), each (n, 3) binary64."""
    n = len(kind)
    sph = kind == 0
    mn = np.empty((n, 3))
    mx = np.empty((n, 3))
    rp = r + 1e-6 + np.abs(r) * 1e-6                     # pack.ts:153
    mn[sph] = a[sph] - rp[sph, None]
    mx[sph] = a[sph] + rp[sph, None]
    tri = ~sph
    mn[tri] = np.minimum(np.minimum(a[tri], b[tri]), c[tri]) - 1e-6
    mx[tri] = np.maximum(np.maximum(a[tri], b[tri]), c[tri]) + 1e-6
    cen = 0.5 * (mn + mx)
    return mn, mx, cen


def _prim_records(kind, a, b, c, r, obj):
    n = len(kind)
    P = np.zeros((max(1, n), PRIM_STRIDE), np.float32)
    if n == 0:
        return P
    P[:n, 0:3] = a
    P[:n, 3] = kind
    P[:n, 7] = obj
    sph = kind == 0
    P[np.flatnonzero(sph), 4] = r[sph]
    P[np.flatnonzero(sph), 11] = 4 * np.pi * r[sph] * r[sph]
    t = np.flatnonzero(~sph)
    if len(t):
        e1 = b[t] - a[t]
        e2 = c[t] - a[t]
        cx = e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1]   # vec3.ts:22
        cy = e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2]
        cz = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
        ss = cx * cx + cy * cy + cz * cz
        ln = np.sqrt(ss)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = np.where(ln > 0, 1 / np.where(ln > 0, ln, 1), 0.0)
        nrm = np.stack([cx * inv, cy * inv, cz * inv], 1)
        nrm[~(ln > 0)] = 0
        P[t, 4:7] = b[t]
        P[t, 8:11] = c[t]
        P[t, 11] = 0.5 * np.sqrt(ss)                      # scene.ts:124-127
        P[t, 12:15] = nrm
    return P


# Primitive count from which the float32 leaf words of the reference layout
# (start*8 + count) stop being exact and the exact words travel beside them.
EXACT_LEAF_LIMIT = 1 << 21


def buildBvh(bmin, bmax, bcen, bins=12, leaf_words=False):
    """buildBvh (bvh.ts:299) on the GPU. Inputs are (n,3) binary64 PrimBounds
    arrays (host numpy or DeviceBuffers). Returns (nodes f32, nodeCount, order u32),
    plus the exact uint32 leaf word of every node when leaf_words is set."""
    A.init()
    n = len(bmin) if not isinstance(bmin, DeviceBuffer) else bmin.shape[0]
    dev = [x if isinstance(x, DeviceBuffer) else
           DeviceBuffer.from_array(np.ascontiguousarray(x, np.float64).reshape(-1, 3))
           for x in (bmin, bmax, bcen)]
    d_nodes = DeviceBuffer.empty((max(1, 2 * n - 1) * NODE_STRIDE,), np.float32)
    d_order = DeviceBuffer.empty((max(1, n),), np.uint32)
    nc = C.c_uint32(0)
    if leaf_words:
        d_words = DeviceBuffer.empty((max(1, 2 * n - 1),), np.uint32)
        check(lib().ct_bvh_build_ex(dev[0].ptr, dev[1].ptr, dev[2].ptr, n, bins, d_nodes.ptr,
                                    C.byref(nc), d_order.ptr, d_words.ptr, None),
              "ct_bvh_build_ex")
    else:
        check(lib().ct_bvh_build(dev[0].ptr, dev[1].ptr, dev[2].ptr, n, bins, d_nodes.ptr,
                                 C.byref(nc), d_order.ptr, None), "ct_bvh_build")
    # only the nodes written come back (the buffer holds room for 2n - 1)
    nodes = d_nodes.numpy(nc.value * NODE_STRIDE)
    order = d_order.numpy(n)
    if leaf_words:
        return nodes, nc.value, order, d_words.numpy(nc.value)
    return nodes, nc.value, order


def packScene(scene, bins=12):
    """packScene (pack.ts:195-223) with the BVH built on the GPU."""
    kind, a, b, c, r, obj = _expand(scene)
    n = len(kind)
    recs = _prim_records(kind, a, b, c, r, obj)
    mn, mx, cen = primBounds(kind, a, b, c, r)
    words = None
    if n >= EXACT_LEAF_LIMIT:
        nodes, nc, order, words = buildBvh(mn, mx, cen, bins, leaf_words=True)
    else:
        nodes, nc, order = buildBvh(mn, mx, cen, bins)
    prims = np.zeros((max(1, n), PRIM_STRIDE), np.float32)
    primObject = np.zeros(max(1, n), np.uint32)
    if n:
        prims[:n] = recs[order]
        primObject[:n] = obj[order]
    objs = scene["objects"]
    mats = np.zeros(max(1, len(objs)) * MAT_STRIDE, np.float32)
    emits = np.zeros(max(1, len(objs)), bool)
    for i, ob in enumerate(objs):
        write_material(mats, i, ob["material"])
        emits[i] = material_emission(ob["material"]) is not None
    lights = np.flatnonzero(emits[primObject[:n]]).astype(np.int32) if n else np.zeros(0, np.int32)
    ps = PackedScene(prims.reshape(-1), n, nodes, nc, mats, len(objs), lights, primObject,
                     scene.get("sky", {"kind": "none"}), order)
    ps.leafWords = words  # layout extension, set from EXACT_LEAF_LIMIT primitives on
    return ps
