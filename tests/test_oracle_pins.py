"""Pins of the oracle's parts that neither the golden image nor bvh.test.ts
fixes: the BVH build's split choices and layout (bvh.ts:111-296), sphere-light
sampling (tracer.ts:61-136), the thin lens (tracer.ts:373-383, camera.ts:62-72)
and the sky gradient (tracer.ts:211-216).

Each part is checked against something other than the oracle's own code:
  * BVH: node arrays derived by hand from bvh.ts's rules
    (tests/golden/bvh_fixtures.json, the derivations are in
    make_bvh_fixtures.py), and a builder written from the definition of the
    binned SAH -- every axis and split costed from scratch on the primitive
    sets themselves, the first strict minimum kept -- on a few hundred small
    inputs with many equal costs;
  * sphere lights: geometry (the cone a sphere subtends, the near root on the
    sphere), the uniform-in-solid-angle law of the cone sample, and the
    closed-form irradiance of a sphere (E = pi L sin^2(a) cos(t) outside it,
    E = pi L inside it) through full renders with and without NEE;
  * the lens: ray origins uniform on the lens disk, every ray of a pixel through
    that pixel's footprint on the focal plane;
  * the sky: the two colours at the zenith and the nadir, affine in between.
Every check must also fail on the oracle built with a planted mistake
(oracle/o64.h, O64_PLANT_*), so that each has teeth.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


# ---------------------------------------------------------------------------
# BVH: hand-derived fixtures

def _fixtures():
    with open(os.path.join(HERE, "golden", "bvh_fixtures.json")) as f:
        return json.load(f)


def check_fixtures(O):
    bad = []
    for fx in _fixtures():
        nodes, nc, order = O.build_bvh(np.array(fx["bmin"]), np.array(fx["bmax"]),
                                       np.array(fx["bcen"]), fx["bins"])
        if (nc != fx["nodeCount"] or list(order) != fx["order"] or
                not np.array_equal(nodes.view(np.uint32), np.array(fx["nodes_u32"], np.uint32))):
            bad.append(fx["name"])
    return bad


# ---------------------------------------------------------------------------
# BVH: a builder from the definition

def _sa(lo, hi):
    dx, dy, dz = hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]
    return 2 * (dx * dy + dy * dz + dz * dx)


def _bucket(c, cmin, scale, bins):
    b = math.floor((c - cmin) * scale)
    return bins - 1 if b >= bins else (0 if b < 0 else b)


def build_by_definition(bmin, bmax, bcen, bins):
    """buildBvh with every SAH split costed from its primitive sets."""
    n = len(bcen)
    order = list(range(n))
    nodes = []

    def box(ids):
        lo = [min(bmin[i][k] for i in ids) for k in range(3)]
        hi = [max(bmax[i][k] for i in ids) for k in range(3)]
        return lo, hi

    def best_split(ids):
        best = None
        for axis in range(3):
            cmin = min(bcen[i][axis] for i in ids)
            cmax = max(bcen[i][axis] for i in ids)
            extent = cmax - cmin
            if not extent > 1e-12:
                continue
            scale = bins / extent
            bk = {i: _bucket(bcen[i][axis], cmin, scale, bins) for i in ids}
            for s in range(bins - 1):
                left = [i for i in ids if bk[i] <= s]
                right = [i for i in ids if bk[i] > s]
                if not left or not right:
                    continue
                cost = len(left) * _sa(*box(left)) + len(right) * _sa(*box(right))
                if best is None or cost < best[0]:
                    best = (cost, axis, s, cmin, scale)
        return best

    def node(start, end):
        idx = len(nodes)
        nodes.append(None)
        ids = order[start:end]
        lo, hi = box(ids)
        nd = np.zeros(8, np.float32)
        for k in range(3):
            nd[k] = np.float32(lo[k] - abs(lo[k]) * 1.2e-7 - 1e-9)
            nd[4 + k] = np.float32(hi[k] + abs(hi[k]) * 1.2e-7 + 1e-9)
        nodes[idx] = nd
        if end - start <= 4:
            nd[3] = idx + 1
            nd[7] = start * 8 + (end - start)
            return
        mid = -1
        b = best_split(ids)
        if b is not None:
            _, axis, s, cmin, scale = b
            lo_i, hi_i = start, end - 1  # the in-place two-pointer partition
            while lo_i <= hi_i:
                if _bucket(bcen[order[lo_i]][axis], cmin, scale, bins) <= s:
                    lo_i += 1
                else:
                    order[lo_i], order[hi_i] = order[hi_i], order[lo_i]
                    hi_i -= 1
            mid = -1 if lo_i in (start, end) else lo_i
        if mid <= start or mid >= end:
            ex, ey, ez = [max(bcen[i][k] for i in ids) - min(bcen[i][k] for i in ids)
                          for k in range(3)]
            axis = 0 if ex >= ey and ex >= ez else (1 if ey >= ez else 2)
            order[start:end] = sorted(order[start:end], key=lambda i: bcen[i][axis])
            mid = (start + end) >> 1
        nd[7] = -1
        node(start, mid)
        node(mid, end)
        nd[3] = len(nodes)

    if n == 0:
        return np.array([0, 0, 0, 1, 0, 0, 0, 0], np.float32), 1, np.zeros(0, np.uint32)
    node(0, n)
    return np.concatenate(nodes), len(nodes), np.array(order, np.uint32)


def _definition_cases():
    rs = np.random.default_rng(2024)
    cases = []
    for k in range(360):
        n = int(rs.integers(5, 17)) if k % 4 else int(rs.integers(17, 41))
        bins = int(rs.choice([2, 3, 4, 12, 16]))
        mode = k % 5
        if mode == 0:    # integer grid: many equal costs and coincident centroids
            cen = rs.integers(0, 4, (n, 3)).astype(np.float64)
            half = rs.choice([0.25, 0.5, 1.0], (n, 1)) * np.ones((1, 3))
        elif mode == 1:  # a plane of the grid (one axis skipped)
            cen = rs.integers(0, 6, (n, 3)).astype(np.float64)
            cen[:, int(rs.integers(0, 3))] = 1.5
            half = np.full((n, 3), 0.5)
        elif mode == 2:  # general position
            cen = rs.uniform(-5, 5, (n, 3))
            half = rs.uniform(0.01, 2, (n, 3))
        elif mode == 3:  # clusters below the 1e-12 extent, among others
            cen = np.repeat(rs.uniform(-3, 3, (3, 3)), (n + 2) // 3, axis=0)[:n]
            cen += rs.integers(0, 4, (n, 3)) * 1e-14
            half = rs.uniform(0.1, 1, (n, 3))
        else:            # symmetric about the x = z diagonal
            m = (n + 1) // 2
            a = rs.integers(0, 5, (m, 3)).astype(np.float64)
            cen = np.concatenate([a, a[:, [2, 1, 0]]])[:n]
            half = np.full((n, 3), 0.5)
        cases.append((cen - half, cen + half, cen, bins))
    return cases


def check_definition(O, cases=None):
    bad = []
    for k, (bmin, bmax, bcen, bins) in enumerate(cases or _definition_cases()):
        dn, dc, do = build_by_definition(bmin.tolist(), bmax.tolist(), bcen.tolist(), bins)
        on, oc, oo = O.build_bvh(bmin.ravel(), bmax.ravel(), bcen.ravel(), bins)
        if oc != dc or not np.array_equal(on.view(np.uint32), dn.view(np.uint32)) or \
                not np.array_equal(oo, do):
            bad.append(k)
    return bad


def test_bvh_matches_hand_derived_fixtures(O):
    assert check_fixtures(O) == []


def test_bvh_matches_the_definition_of_binned_sah(O):
    assert check_definition(O) == []


def test_definition_builder_against_fixtures():
    """The definition builder reproduces the hand derivations too."""
    for fx in _fixtures():
        b = np.array(fx["bmin"]).reshape(-1, 3).tolist()
        t = np.array(fx["bmax"]).reshape(-1, 3).tolist()
        c = np.array(fx["bcen"]).reshape(-1, 3).tolist()
        nodes, nc, order = build_by_definition(b, t, c, fx["bins"])
        assert nc == fx["nodeCount"] and list(order) == fx["order"], fx["name"]
        assert np.array_equal(nodes.view(np.uint32), np.array(fx["nodes_u32"], np.uint32))


# ---------------------------------------------------------------------------
# Sphere lights

def _lambert(a):
    return {"type": "lambertian", "albedo": (a, a, a)}


def _emitter(le):
    return {"type": "emissive", "color": (1.0, 1.0, 1.0), "strength": le}


def _quad(o, u, v):
    return {"kind": "quad", "origin": o, "u": u, "v": v}


def _sphere(c, r):
    return {"kind": "sphere", "center": c, "radius": r}


CAM = {"target": (0.0, 0.0, 0.0), "yaw": 0.0, "pitch": 0.4, "distance": 10.0, "fovDeg": 1.0,
       "lensRadius": 0.0, "focusDistance": 10.0}


def _light_over_plane(O):
    """A diffuse plane (albedo 0.5) under an emissive sphere (radius 1,
    radiance 4) centred 3 above the origin."""
    return O.pack_scene({"objects": [
        {"name": "floor", "shape": _quad((-50.0, 0.0, -50.0), (100.0, 0.0, 0.0),
                                         (0.0, 0.0, 100.0)), "material": _lambert(0.5)},
        {"name": "lamp", "shape": _sphere((0.0, 3.0, 0.0), 1.0), "material": _emitter(4.0)}],
        "camera": CAM, "sky": {"kind": "none"}})


def _ks_uniform(x):
    """Kolmogorov-Smirnov distance of samples in [0, 1] from U(0, 1)."""
    x = np.sort(np.asarray(x))
    n = len(x)
    i = np.arange(1, n + 1)
    return max(np.max(i / n - x), np.max(x - (i - 1) / n))


def check_cone(O):
    """Light samples from points outside the sphere: every direction inside the
    subtended cone (sin a = r/d), cos(angle) uniform on [cos a, 1] and the
    azimuth uniform (uniform in solid angle), the pdf one over the cone's solid
    angle 2 pi (1 - cos a), and the distance the near root on the sphere."""
    ps = _light_over_plane(O)
    c = np.array([0.0, 3.0, 0.0])
    r = 1.0
    bad = []
    for p in (np.array([0.0, 0.0, 0.0]), np.array([2.0, 0.0, -1.5]), np.array([0.3, 1.4, 0.2])):
        axis = c - p
        d = np.linalg.norm(axis)
        axis /= d
        cos_a = math.sqrt(1 - (r / d) ** 2)
        solid = 2 * math.pi * (1 - cos_a)
        ref = np.array([1.0, 0.0, 0.0]) if abs(axis[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e1 = np.cross(axis, ref)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(axis, e1)
        u, phi = [], []
        for k in range(4000):
            ok, out, _ = O.sample_light(ps, p, O.seed_for(k, 17))
            if not ok:
                continue
            w, dist, pdf = out[0:3], out[3], out[4]
            cw = float(np.dot(w, axis))
            if cw < cos_a - 1e-12 or abs(np.linalg.norm(w) - 1) > 1e-12:
                bad.append(f"direction outside the cone from {p}")
                break
            if abs(pdf * solid - 1) > 1e-12:
                bad.append(f"pdf {pdf} is not 1/(2 pi (1 - cos a)) = {1 / solid}")
                break
            q = p + dist * w
            if abs(np.linalg.norm(q - c) - r) > 1e-9 or dist > d:
                bad.append("distance is not the near root on the sphere")
                break
            if abs(O.light_pdf(ps, p, q, (q - c) / r, int(ps.lights[0])) * solid - 1) > 1e-12:
                bad.append("lightPdfScalar differs from the sample's pdf")
                break
            u.append((1 - cw) / (1 - cos_a))
            phi.append((math.atan2(np.dot(w, e2), np.dot(w, e1)) / (2 * math.pi)) % 1.0)
        if bad:
            break
        if len(u) < 3900:
            bad.append(f"only {len(u)} of 4000 samples accepted")
        elif _ks_uniform(u) > 0.04 or _ks_uniform(phi) > 0.04:
            bad.append(f"cone sample not uniform in solid angle from {p}: "
                       f"KS {_ks_uniform(u):.3f} {_ks_uniform(phi):.3f}")
    return bad


def _mean_and_sigma(O, ps, frame, w, h, passes, **kw):
    vals = np.concatenate([O.pass_radiance(ps, frame, w, h, s, **kw) for s in range(passes)])
    v = vals[:, 0]
    return v.mean(), v.std() / math.sqrt(len(v))


def check_sphere_irradiance(O):
    """Outgoing radiance of the diffuse plane under the sphere light: albedo/pi
    times E = pi L sin^2(a) cos(t), a the cone's half angle and t the angle
    between the normal and the centre direction, averaged over the pixels'
    points; estimated with NEE + MIS and with BSDF sampling alone."""
    ps = _light_over_plane(O)
    w = h = 4
    frame = O.camera_frame(CAM, 1.0)
    expect = []
    for y in range(h):
        for x in range(w):
            o, d = O.generate_ray(frame, (x + 0.5) / w, 1 - (y + 0.5) / h)
            t = -o[1] / d[1]
            q = o + t * d
            v = np.array([0.0, 3.0, 0.0]) - q
            dist = np.linalg.norm(v)
            expect.append(0.5 / math.pi * math.pi * 4.0 * (1.0 / dist) ** 2 * (v[1] / dist))
    expect = float(np.mean(expect))
    bad = []
    for nee, passes in ((True, 4096), (False, 4096)):
        m, sig = _mean_and_sigma(O, ps, frame, w, h, passes, max_bounces=1,
                                 russian_roulette=False, nee=nee)
        if abs(m - expect) > 5 * sig + 2e-4 * expect:
            bad.append(f"nee={nee}: {m:.5f} +- {sig:.5f}, closed form {expect:.5f}")
    return bad


def check_inside_sphere(O):
    """A diffuse plane inside an emissive sphere of radiance L sees L over its
    whole upper hemisphere: outgoing radiance albedo * L exactly (NEE draws
    uniform points on the sphere, the inside case of sampleLightInto)."""
    ps = O.pack_scene({"objects": [
        {"name": "floor", "shape": _quad((-30.0, 0.0, -30.0), (60.0, 0.0, 0.0),
                                         (0.0, 0.0, 60.0)), "material": _lambert(0.5)},
        {"name": "dome", "shape": _sphere((0.0, 0.0, 0.0), 10.0), "material": _emitter(1.0)}],
        "camera": CAM, "sky": {"kind": "none"}})
    cam = dict(CAM, distance=5.0, fovDeg=20.0)
    frame = O.camera_frame(cam, 1.0)
    bad = []
    for nee in (True, False):
        m, sig = _mean_and_sigma(O, ps, frame, 4, 4, 512, max_bounces=1,
                                 russian_roulette=False, nee=nee)
        if abs(m - 0.5) > 5 * sig + 1e-4:
            bad.append(f"nee={nee}: {m:.5f} +- {sig:.5f}, closed form 0.5")
    return bad


def test_sphere_light_cone_sampling(O):
    assert check_cone(O) == []


def test_sphere_light_irradiance_closed_form(O):
    assert check_sphere_irradiance(O) == []


def test_inside_sphere_light_closed_form(O):
    assert check_inside_sphere(O) == []


# ---------------------------------------------------------------------------
# Thin lens

def check_lens(O):
    """Ray origins of one pixel uniform on the lens disk (radius^2 and angle
    uniform), in the lens plane, and every ray through that pixel's footprint
    on the focal plane; a pinhole camera's rays all start at its origin."""
    cam = {"target": (0.0, 1.0, 0.0), "yaw": 0.7, "pitch": 0.2, "distance": 9.0,
           "fovDeg": 40.0, "lensRadius": 0.8, "focusDistance": 7.0}
    w, h = 8, 6
    frame = O.camera_frame(cam, w / h)
    org, right, up, fwd = frame[0:3], frame[3:6], frame[6:9], frame[9:12]
    tan_h, aspect, lr, fd = frame[12], frame[13], frame[14], frame[15]
    bad = []
    for (x, y) in ((3, 2), (0, 5), (7, 0)):
        r2, ang = [], []
        u0 = (2 * x / w - 1) * tan_h * aspect
        u1 = (2 * (x + 1) / w - 1) * tan_h * aspect
        v0 = (2 * (1 - (y + 1) / h) - 1) * tan_h
        v1 = (2 * (1 - y / h) - 1) * tan_h
        for k in range(3000):
            o, d = O.camera_sample(frame, w, h, x, y, k)
            off = o - org
            if abs(np.dot(off, fwd)) > 1e-12 or abs(np.linalg.norm(d) - 1) > 1e-12:
                bad.append("origin off the lens plane or direction not unit")
                break
            a, b = np.dot(off, right) / lr, np.dot(off, up) / lr
            if a * a + b * b > 1 + 1e-12:
                bad.append("origin outside the lens disk")
                break
            r2.append(a * a + b * b)
            ang.append((math.atan2(b, a) / (2 * math.pi)) % 1.0)
            t = fd / np.dot(d, fwd)  # the origin is in the lens plane
            q = o + t * d - org
            qu, qv = np.dot(q, right) / fd, np.dot(q, up) / fd
            if not (u0 - 1e-9 <= qu <= u1 + 1e-9 and v0 - 1e-9 <= qv <= v1 + 1e-9):
                bad.append(f"pixel ({x},{y}) ray misses its footprint on the focal plane")
                break
        if bad:
            break
        if _ks_uniform(r2) > 0.045 or _ks_uniform(ang) > 0.045:
            bad.append(f"lens samples not uniform on the disk: KS {_ks_uniform(r2):.3f} "
                       f"{_ks_uniform(ang):.3f}")
    pin = O.camera_frame(dict(cam, lensRadius=0.0), w / h)
    for k in range(50):
        o, _ = O.camera_sample(pin, w, h, 2, 3, k)
        if not np.array_equal(o, pin[0:3]):
            bad.append("a pinhole ray does not start at the camera origin")
            break
    return bad


def test_thin_lens(O):
    assert check_lens(O) == []


# ---------------------------------------------------------------------------
# Sky gradient

def check_sky(O):
    """Radiance of a miss: the zenith colour times strength straight up, the
    horizon colour times strength straight down, affine in the direction's y
    in between and independent of x and z."""
    zen, hor, k = (0.2, 0.4, 0.9), (0.9, 0.7, 0.5), 1.7
    ps = O.pack_scene({"objects": [], "camera": CAM,
                       "sky": {"kind": "gradient", "zenith": zen, "horizon": hor, "strength": k}})
    L = lambda d: O.radiance(ps, (0.0, 0.0, 0.0), np.asarray(d, float) / np.linalg.norm(d), 5)
    bad = []
    if not np.allclose(L((0, 1, 0)), np.array(zen) * k, rtol=1e-12, atol=0):
        bad.append("straight up is not the zenith colour")
    if not np.allclose(L((0, -1, 0)), np.array(hor) * k, rtol=1e-12, atol=0):
        bad.append("straight down is not the horizon colour")
    top, bottom = L((0, 1, 0)), L((0, -1, 0))
    for dy in (-0.8, -0.3, 0.0, 0.45, 0.9):
        s = math.sqrt(1 - dy * dy)
        for ang in (0.0, 2.0, 4.0):
            d = (s * math.cos(ang), dy, s * math.sin(ang))
            f = (dy + 1) / 2
            if not np.allclose(L(d), bottom + (top - bottom) * f, rtol=1e-12, atol=1e-15):
                bad.append(f"not affine in y at dy={dy}")
                break
    return bad


def test_sky_gradient(O):
    assert check_sky(O) == []


def test_renders_use_the_camera_samples(O):
    """A pass over an empty scene with a sky: each path's radiance is the sky
    along the ray camera_sample returns for that pixel and pass."""
    ps = O.pack_scene({"objects": [], "camera": CAM,
                       "sky": {"kind": "gradient", "zenith": (0.1, 0.2, 0.3),
                               "horizon": (1.0, 0.5, 0.25), "strength": 2.0}})
    cam = dict(CAM, lensRadius=0.5, fovDeg=60.0)
    frame = O.camera_frame(cam, 1.5)
    rad = O.pass_radiance(ps, frame, 6, 4, 9)
    for y in range(4):
        for x in range(6):
            o, d = O.camera_sample(frame, 6, 4, x, y, 9)
            assert np.array_equal(rad[y * 6 + x], O.radiance(ps, o, d, 0, max_bounces=0))


# ---------------------------------------------------------------------------
# Primitive records, PrimBounds and the camera frame from geometry

def test_prim_records_and_bounds_from_geometry(O):
    """pack.ts:131-192: a triangle's record holds its vertices, the unit normal
    of its plane and half the cross product's length as area; a sphere holds
    its centre, radius and 4 pi r^2.  Bounds: the vertices' box widened by 1e-6
    per side, a sphere's cube of half side r + 1e-6 + |r| 1e-6, centroids the
    boxes' centres (bvh.ts:46-58)."""
    rs = np.random.default_rng(5)
    tris = rs.integers(-8, 9, (40, 3, 3)).astype(np.float64) / 4
    objects = [{"name": f"t{i}", "shape": {"kind": "triangle", "a": tuple(t[0]),
                                            "b": tuple(t[1]), "c": tuple(t[2])},
                "material": _lambert(0.5)} for i, t in enumerate(tris)]
    objects += [{"name": "s", "shape": _sphere((1.5, -2.25, 0.5), 0.75),
                 "material": _lambert(0.5)}]
    sc = {"objects": objects, "camera": CAM, "sky": {"kind": "none"}}
    ps = O.pack_scene(sc)
    kind, geom, _ = O.expand_scene(sc)
    bmin, bmax, bcen = [x.reshape(-1, 3) for x in O.prim_bounds(kind, geom)]
    for src in range(len(objects)):
        dst = int(np.flatnonzero(ps.order == src)[0])
        rec = ps.prims[16 * dst: 16 * dst + 16].astype(np.float64)
        if src < len(tris):
            t = tris[src]
            assert rec[3] == 1  # PRIM_TRIANGLE (pack.ts:39)
            assert np.array_equal(rec[[0, 1, 2, 4, 5, 6, 8, 9, 10]], t.ravel())
            cr = np.cross(t[1] - t[0], t[2] - t[0])
            ln = np.linalg.norm(cr)
            if ln > 0:
                assert np.allclose(rec[12:15], cr / ln, rtol=0, atol=1e-7)
            assert abs(rec[11] - 0.5 * ln) <= 1e-6 * max(1.0, ln)
            assert np.array_equal(bmin[src], t.min(axis=0) - 1e-6)
            assert np.array_equal(bmax[src], t.max(axis=0) + 1e-6)
        else:
            assert rec[3] == 0  # PRIM_SPHERE (pack.ts:38)
            assert np.array_equal(rec[[0, 1, 2, 4]], [1.5, -2.25, 0.5, 0.75])
            assert abs(rec[11] - 4 * math.pi * 0.75 ** 2) <= 1e-6
            pad = 0.75 + 1e-6 + 0.75 * 1e-6
            assert np.allclose(bmin[src], np.array([1.5, -2.25, 0.5]) - pad, rtol=0, atol=1e-15)
            assert np.allclose(bmax[src], np.array([1.5, -2.25, 0.5]) + pad, rtol=0, atol=1e-15)
        assert np.allclose(bcen[src], 0.5 * (bmin[src] + bmax[src]), rtol=0, atol=1e-15)


def test_camera_frame_from_geometry(O):
    """camera.ts:25-51: an orthonormal right-handed basis with forward pointing
    from the orbit position to the target, right horizontal, up = right x
    forward; tanHalfFov = tan(fov/2); the focus distance kept >= 1e-3."""
    for yaw, pitch, dist, fov in ((0.0, 0.0, 5.0, 40.0), (0.7, 0.3, 9.0, 60.0),
                                  (-2.0, -1.2, 3.0, 20.0), (1.0, 1.5, 2.0, 90.0)):
        cam = {"target": (1.0, -2.0, 0.5), "yaw": yaw, "pitch": pitch, "distance": dist,
               "fovDeg": fov, "lensRadius": 0.1, "focusDistance": 0.0}
        f = O.camera_frame(cam, 1.25)
        org, right, up, fwd = f[0:3], f[3:6], f[6:9], f[9:12]
        pos = np.array(cam["target"]) + dist * np.array(
            [math.cos(pitch) * math.sin(yaw), math.sin(pitch), math.cos(pitch) * math.cos(yaw)])
        assert np.allclose(org, pos, atol=1e-12)
        assert np.allclose(fwd, (np.array(cam["target"]) - pos) / dist, atol=1e-12)
        for v in (right, up, fwd):
            assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert abs(np.dot(right, up)) < 1e-12 and abs(np.dot(right, fwd)) < 1e-12
        assert abs(np.dot(up, fwd)) < 1e-12 and abs(right[1]) < 1e-12 and up[1] > 0
        assert np.allclose(np.cross(right, fwd), up, atol=1e-12)
        assert abs(f[12] - math.tan(math.radians(fov) / 2)) < 1e-15
        assert f[13] == 1.25 and f[14] == 0.1 and f[15] == 1e-3


# ---------------------------------------------------------------------------
# Planted mistakes: each pin above must catch its mistake

PLANT_CHECKS = {
    "sah_le": ["fixtures"], "axis_reversed": ["fixtures"], "nudge_inward": ["fixtures"],
    "partition_lt": ["fixtures"], "median_unstable": ["fixtures"],
    "median_axis_last": ["fixtures"], "sweep_shift": ["definition"], "leaf3": ["definition"],
    "cone_no_sqrt": ["cone", "irradiance"], "cone_angle_uniform": ["cone"],
    "cone_pdf_2x": ["cone", "irradiance"], "inside_half_area": ["inside"],
    "lens_linear_r": ["lens"], "lens_no_refocus": ["lens"], "sky_flip": ["sky"],
}
CHECKS = {"fixtures": check_fixtures, "definition": check_definition, "cone": check_cone,
          "irradiance": check_sphere_irradiance, "inside": check_inside_sphere,
          "lens": check_lens, "sky": check_sky}


@pytest.fixture(scope="module")
def planted(O, tmp_path_factory):
    return O.build_planted(sorted(PLANT_CHECKS), str(tmp_path_factory.mktemp("plants")))


@pytest.mark.parametrize("name", sorted(PLANT_CHECKS))
def test_planted_mistake_is_caught(O, planted, name):
    with O.using_library(planted[name]):
        for check in PLANT_CHECKS[name]:
            assert CHECKS[check](O), f"{check} does not catch the planted mistake {name}"
