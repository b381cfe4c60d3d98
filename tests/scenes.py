"""Scene fixtures used by the tests, written in the reference's plain-object scene
format (src/tracer/scene.ts:5-74) as Python dicts.

Random layouts draw from the reference RNG stream (rng.ts) in exactly the order
the reference generators do, so the packed arrays equal the reference's.
"""
import math

import numpy as np


def rgb(r, g, b):
    return (r, g, b)


def lambert(albedo, checker=None, emission=None):
    m = {"type": "lambertian", "albedo": albedo}
    if checker is not None:
        m["checker"] = checker
    if emission is not None:
        m["emission"] = emission
    return m


def metal(albedo, roughness):
    return {"type": "metal", "albedo": albedo, "roughness": roughness}


def glass(ior=1.5, tint=None):
    m = {"type": "dielectric", "ior": ior}
    if tint is not None:
        m["tint"] = tint
    return m


def light(color, strength):
    return {"type": "emissive", "color": color, "strength": strength}


def quad(o, u, v):
    return {"kind": "quad", "origin": o, "u": u, "v": v}


def sphere(c, r):
    return {"kind": "sphere", "center": c, "radius": r}


def obj(name, shape, material):
    return {"name": name, "shape": shape, "material": material}


def camera_at(eye, target, fov_deg, lens_radius, focus_distance=None):
    """cameraAt (scenes.ts:26-41)."""
    dx, dy, dz = eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]
    distance = math.sqrt(dx * dx + dy * dy + dz * dz)
    return {"target": tuple(target), "yaw": math.atan2(dx, dz), "pitch": math.asin(dy / distance),
            "distance": distance, "fovDeg": fov_deg, "lensRadius": lens_radius,
            "focusDistance": distance if focus_distance is None else focus_distance}


def cornell_box(light_color=(1, 0.86, 0.7), light_strength=12, metal_roughness=0.02,
                checker_back_wall=None):
    """buildCornellBox (scenes.ts:45-67). The keyword arguments give the
    256x256 variant (radiance-15 light, roughness-0.1 metal, checker back wall)."""
    white = lambert(rgb(0.73, 0.73, 0.73))
    red = lambert(rgb(0.65, 0.05, 0.05))
    green = lambert(rgb(0.12, 0.45, 0.15))
    back = white if checker_back_wall is None else lambert(rgb(0.73, 0.73, 0.73),
                                                          checker=checker_back_wall)
    W = 0.75
    objects = [
        obj("Floor", quad((-W, 0, -0.5), (2 * W, 0, 0), (0, 0, 1)), white),
        obj("Ceiling", quad((-W, 1, -0.5), (2 * W, 0, 0), (0, 0, 1)), white),
        obj("Back wall", quad((-W, 0, -0.5), (2 * W, 0, 0), (0, 1, 0)), back),
        obj("Red wall", quad((-W, 0, -0.5), (0, 0, 1), (0, 1, 0)), red),
        obj("Green wall", quad((W, 0, -0.5), (0, 0, 1), (0, 1, 0)), green),
        obj("Ceiling light", quad((-0.2, 0.998, -0.15), (0.4, 0, 0), (0, 0, 0.3)),
            light(light_color, light_strength)),
        obj("Metal sphere", sphere((-0.33, 0.24, -0.12), 0.24), metal(rgb(0.9, 0.9, 0.92),
                                                                       metal_roughness)),
        obj("Glass sphere", sphere((0.3, 0.2, 0.17), 0.2), glass(1.5)),
    ]
    return {"name": "Cornell Box", "objects": objects,
            "camera": camera_at((0, 0.5, 1.6), (0, 0.5, 0), 32, 0, 1.6),
            "sky": {"kind": "none"}}


def cornell_c1():
    """The 256x256 / 16 spp Cornell variant: 2-triangle light of radiance 15,
    metal roughness 0.1, checker-modulated back wall."""
    return cornell_box(light_color=(1, 1, 1), light_strength=15, metal_roughness=0.1,
                       checker_back_wall={"color2": rgb(0.25, 0.25, 0.25), "scale": 8})


def random_scene(rng, spheres, triangles):
    """randomScene (bvh.test.ts:9-42); rng is an object with range(lo, hi)."""
    objects = []
    grey = lambert(rgb(0.5, 0.5, 0.5))
    for i in range(spheres):
        c = (rng.range(-10, 10), rng.range(-10, 10), rng.range(-10, 10))
        r = rng.range(0.05, 0.6)
        objects.append(obj(f"s{i}", sphere(c, r), grey))
    for i in range(triangles):
        a = (rng.range(-10, 10), rng.range(-10, 10), rng.range(-10, 10))
        b = (a[0] + rng.range(-1, 1), a[1] + rng.range(-1, 1), a[2] + rng.range(-1, 1))
        c = (a[0] + rng.range(-1, 1), a[1] + rng.range(-1, 1), a[2] + rng.range(-1, 1))
        objects.append(obj(f"t{i}", {"kind": "triangle", "a": a, "b": b, "c": c}, grey))
    return {"name": "random", "objects": objects,
            "camera": {"target": (0, 0, 0), "yaw": 0, "pitch": 0, "distance": 30, "fovDeg": 40,
                       "lensRadius": 0, "focusDistance": 30},
            "sky": {"kind": "none"}}


def furnace():
    """FURNACE (furnace.test.ts:13-25)."""
    return {"name": "furnace",
            "objects": [obj("shell", sphere((0, 0, 0), 5),
                            lambert(rgb(0.8, 0.8, 0.8), emission=rgb(1, 1, 1)))],
            "camera": {"target": (0, 0, -1), "yaw": 0, "pitch": 0, "distance": 1, "fovDeg": 60,
                       "lensRadius": 0, "focusDistance": 1},
            "sky": {"kind": "none"}}


def coincident_spheres(n=40):
    """bvh.test.ts:153-168: n identical spheres (median-split fallback)."""
    return {"name": "coincident",
            "objects": [obj("c", sphere((1, 2, 3), 0.5), lambert(rgb(1, 1, 1))) for _ in range(n)],
            "camera": {"target": (0, 0, 0), "yaw": 0, "pitch": 0, "distance": 30, "fovDeg": 40,
                       "lensRadius": 0, "focusDistance": 30},
            "sky": {"kind": "none"}}


def glass_spheres(rng):
    """buildGlassSpheres (scenes.ts:72-113): checker ground, lens, gradient sky."""
    objects = [
        obj("Checker ground", quad((-40, 0, -40), (80, 0, 0), (0, 0, 80)),
            lambert(rgb(0.82, 0.8, 0.76), checker={"color2": rgb(0.18, 0.17, 0.16), "scale": 1})),
        obj("Big glass", sphere((0, 1, 0), 1), glass(1.5)),
        obj("Big diffuse", sphere((-4, 1, 0), 1), lambert(rgb(0.4, 0.2, 0.1))),
        obj("Big metal", sphere((4, 1, 0), 1), metal(rgb(0.7, 0.6, 0.5), 0)),
    ]
    bigs = [(0, 1, 0), (-4, 1, 0), (4, 1, 0)]
    n = 0
    for a in range(-6, 6):
        if n >= 60:
            break
        for b in range(-4, 5):
            if n >= 60:
                break
            center = (a * 1.15 + 0.7 * rng.next(), 0.2, b * 1.15 + 0.7 * rng.next())
            if any(math.hypot(c[0] - center[0], c[2] - center[2]) < 1.4 for c in bigs):
                continue
            pick = rng.next()
            if pick < 0.62:
                m = lambert(rgb(rng.next() * rng.next(), rng.next() * rng.next(),
                                rng.next() * rng.next()))
            elif pick < 0.85:
                m = metal(rgb(0.5 + 0.5 * rng.next(), 0.5 + 0.5 * rng.next(),
                              0.5 + 0.5 * rng.next()), 0.3 * rng.next())
            else:
                m = glass(1.5)
            objects.append(obj(f"Sphere {n + 1}", sphere(center, 0.2), m))
            n += 1
    return {"name": "Glass Spheres on Checkerboard", "objects": objects,
            "camera": camera_at((12, 2.2, 3.2), (0, 0.55, 0), 22, 0.06, 10.5),
            "sky": {"kind": "gradient", "zenith": rgb(0.5, 0.7, 1), "horizon": rgb(1, 1, 1),
                    "strength": 1}}


def night_emissives():
    """buildNightEmissives (scenes.ts:118-142): sphere lights only."""
    glows = [((-1.6, 0.35, -0.6), 0.35, rgb(1, 0.35, 0.1), 6),
             ((0.2, 0.25, 0.9), 0.25, rgb(0.2, 0.55, 1), 7),
             ((1.7, 0.45, -1.1), 0.45, rgb(0.6, 1, 0.3), 4),
             ((-0.4, 0.15, -1.9), 0.15, rgb(1, 0.2, 0.6), 10),
             ((1.1, 0.18, 1.6), 0.18, rgb(1, 0.85, 0.4), 8),
             ((-2.6, 0.2, 1.4), 0.2, rgb(0.5, 0.3, 1), 7)]
    objects = [
        obj("Dark ground", quad((-30, 0, -30), (60, 0, 0), (0, 0, 60)),
            lambert(rgb(0.09, 0.085, 0.08))),
        obj("Chrome sphere", sphere((0.4, 0.55, -0.7), 0.55), metal(rgb(0.95, 0.95, 0.95), 0.0)),
        obj("Brushed sphere", sphere((-1.0, 0.4, 1.1), 0.4), metal(rgb(0.9, 0.7, 0.5), 0.25)),
        obj("Glass sphere", sphere((1.9, 0.4, 0.6), 0.4), glass(1.5)),
        obj("Matte sphere", sphere((-2.2, 0.5, -1.5), 0.5), lambert(rgb(0.7, 0.7, 0.72))),
    ]
    for i, (p, r, c, s) in enumerate(glows):
        objects.append(obj(f"Glow {i + 1}", sphere(p, r), light(c, s)))
    return {"name": "Night Emissives", "objects": objects,
            "camera": camera_at((4.2, 1.6, 4.6), (0, 0.35, 0), 30, 0.02, 6.2),
            "sky": {"kind": "none"}}


def js_hypot(*xs):
    """Math.hypot as V8 evaluates it: scale by the largest magnitude, then a
    Kahan-compensated sum of squares (bit-exact against the engine's outputs in
    tests/golden/jsmath_vectors.json)."""
    a = [abs(x) for x in xs]
    m = max(a) if a else 0.0
    if m == math.inf:
        return math.inf
    if m == 0:
        return 0.0
    s = 0.0
    comp = 0.0
    for x in a:
        n = x / m
        summand = n * n - comp
        prelim = s + summand
        comp = (prelim - s) - summand
        s = prelim
    return math.sqrt(s) * m


def ray_set_bvh_test(rng, n, lo=-14, hi=14, tlo=-8, thi=8):
    """bvh.test.ts:94-99: origins in [lo,hi]^3 aimed at [tlo,thi]^3, normalised by hypot."""
    d = rng.take(6 * n).reshape(n, 6)
    o = lo + (hi - lo) * d[:, 0:3]
    tgt = tlo + (thi - tlo) * d[:, 3:6]
    dirv = tgt - o
    ln = np.array([js_hypot(*row) for row in dirv])
    return np.concatenate([o, dirv / ln[:, None]], axis=1)


def tie_rays(ps, seed, per):
    """Rays aimed at the vertices, edge midpoints and an edge quarter point of
    every triangle: where meshes and quads share edges and vertices, several
    primitives are hit at bit-identical t."""
    rs = np.random.default_rng(seed)
    prims = ps.prims.reshape(-1, 16)
    rays = []
    for p in np.flatnonzero(prims[:, 3] == 1):
        a, b, c = (prims[p, k:k + 3].astype(np.float64) for k in (0, 4, 8))
        for tgt in (a, b, c, 0.5 * (a + b), 0.5 * (b + c), 0.5 * (a + c), a + 0.25 * (b - a)):
            for _ in range(per):
                o = tgt + rs.uniform(-6, 6, 3)
                d = tgt - o
                rays.append(np.concatenate([o, d / np.linalg.norm(d)]))
    return np.array(rays)
