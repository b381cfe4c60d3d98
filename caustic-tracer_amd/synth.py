"""Synthetic workloads for the tracer, in the reference's scene format.

All coordinates are rounded to float32 before packing, so the FP64 bounds the
packer derives from the scene equal those derived from the packed records.
"""
import math

import numpy as np


def _f32(a):
    return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def _random_frames(rng, n):
    """Uniform unit normals on the sphere and a uniform in-plane angle."""
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    r = np.sqrt(np.maximum(0, 1 - z * z))
    nrm = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    helper = np.where(np.abs(nrm[:, :1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    t = np.cross(helper, nrm)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(nrm, t)
    ang = rng.uniform(0, 2 * math.pi, n)
    e0 = np.cos(ang)[:, None] * t + np.sin(ang)[:, None] * b
    e1 = np.cos(ang + 2 * math.pi / 3)[:, None] * t + np.sin(ang + 2 * math.pi / 3)[:, None] * b
    e2 = np.cos(ang + 4 * math.pi / 3)[:, None] * t + np.sin(ang + 4 * math.pi / 3)[:, None] * b
    return e0, e1, e2


def equilateral_soup(rng, n, extent=10.0, mu=math.log(0.05), sigma=0.8):
    """n equilateral triangles: centroids uniform in [-extent, extent]^3, edge
    length lognormal(mu, sigma), orientation uniform. Returns (n, 3, 3)."""
    c = rng.uniform(-extent, extent, (n, 3))
    s = rng.lognormal(mu, sigma, n)
    R = s / math.sqrt(3)  # circumradius
    e0, e1, e2 = _random_frames(rng, n)
    tri = np.stack([c + R[:, None] * e0, c + R[:, None] * e1, c + R[:, None] * e2], 1)
    return _f32(tri)


def soup_scene(n_tris=1_000_000, n_lights=64, n_materials=1000, seed=2024, light_edge=0.5,
               light_radiance=20.0):
    """Triangle soup: n_tris triangles spread over n_materials mesh objects whose
    materials are drawn 70% Lambertian / 20% metal / 10% dielectric, plus one
    emissive object of n_lights triangles, gradient sky, camera outside the cloud."""
    rng = np.random.default_rng(seed)
    tris = equilateral_soup(rng, n_tris)
    which = rng.integers(0, n_materials, n_tris)
    kinds = rng.choice(3, size=n_materials, p=[0.7, 0.2, 0.1])
    objects = []
    order = np.argsort(which, kind="stable")
    counts = np.bincount(which, minlength=n_materials)
    offs = np.concatenate([[0], np.cumsum(counts)])
    for m in range(n_materials):
        sel = tris[order[offs[m]:offs[m + 1]]]
        if kinds[m] == 0:
            mat = {"type": "lambertian",
                   "albedo": tuple(float(x) for x in _f32(rng.uniform(0.2, 0.9, 3)))}
        elif kinds[m] == 1:
            mat = {"type": "metal", "albedo": tuple(float(x) for x in _f32(rng.uniform(0.5, 1.0, 3))),
                   "roughness": float(_f32(rng.uniform(0, 0.3)))}
        else:
            mat = {"type": "dielectric", "ior": 1.5}
        objects.append({"name": f"mesh{m}", "material": mat,
                        "shape": {"kind": "mesh", "vertices": sel.reshape(-1, 3),
                                  "indices": np.arange(3 * len(sel))}})
    lt = equilateral_soup(rng, n_lights, mu=math.log(light_edge), sigma=0.0)
    objects.append({"name": "lights", "material": {"type": "emissive", "color": (1.0, 1.0, 1.0),
                                                   "strength": light_radiance},
                    "shape": {"kind": "mesh", "vertices": lt.reshape(-1, 3),
                              "indices": np.arange(3 * n_lights)}})
    eye = (0.0, 0.0, 24.0)
    dist = 24.0
    camera = {"target": (0.0, 0.0, 0.0), "yaw": math.atan2(eye[0], eye[2]),
              "pitch": math.asin(eye[1] / dist), "distance": dist, "fovDeg": 45.0,
              "lensRadius": 0.0, "focusDistance": dist}
    return {"name": f"soup{n_tris}", "objects": objects, "camera": camera,
            "sky": {"kind": "gradient", "zenith": (0.5, 0.7, 1.0), "horizon": (1.0, 1.0, 1.0),
                    "strength": 0.5}}


def triangle_counts(n, seed, lo, hi, median=20_000, sigma=1.0):
    """n triangle counts, lognormal around `median` (log-sd `sigma`), clipped to [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.lognormal(math.log(median), sigma, n)), lo, hi).astype(np.int64)


def family_scene(i, n_tris):
    """Scene i of a family: the soup generator with seed i, materials scaled to
    the triangle count (about 100 triangles per mesh object, at most 1000)."""
    return soup_scene(int(n_tris), n_lights=64, n_materials=int(min(1000, max(8, n_tris // 100))),
                      seed=int(i))


def orbit_camera(rng, dist_lo=20.0, dist_hi=30.0):
    """Camera on a sphere around the soup's centre, looking at it."""
    yaw = float(rng.uniform(-math.pi, math.pi))
    pitch = float(rng.uniform(-0.6, 0.6))
    dist = float(rng.uniform(dist_lo, dist_hi))
    return {"target": (0.0, 0.0, 0.0), "yaw": yaw, "pitch": pitch, "distance": dist,
            "fovDeg": 45.0, "lensRadius": 0.0, "focusDistance": dist}


def batch_workload(n_scenes=64, seed=4, lo=1_000, hi=500_000):
    """Batching form: n_scenes scenes with lognormal triangle counts in [lo, hi]
    (median 20k), scene i generated from seed i; per scene one job with a
    resolution drawn from {128, 256, 512, 768}^2, spp uniform in 1..32 and
    maxBounces 8.  Returns (scenes, jobs) with jobs as dicts of width, height,
    spp, seed, maxBounces and the index of their scene."""
    counts = triangle_counts(n_scenes, seed, lo, hi)
    rng = np.random.default_rng(seed + 1)
    scenes, jobs = [], []
    for i, n in enumerate(counts):
        scenes.append(family_scene(i, n))
        side = int(rng.choice([128, 256, 512, 768]))
        jobs.append({"sceneIndex": i, "width": side, "height": side,
                     "spp": int(rng.integers(1, 33)), "seed": i + 1, "maxBounces": 8,
                     "russianRoulette": True, "nee": True})
    return scenes, jobs


def request_stream(n_scenes=128, rate=200.0, duration=60.0, seed=5, res_lo=128, res_hi=1024,
                   spp_hi=64, depth_lo=4, depth_hi=12):
    """Serving form: Poisson arrivals at `rate` per second for `duration` seconds;
    each request names one of n_scenes resident scenes, a random orbit camera,
    a square resolution that is a multiple of 128 in [res_lo, res_hi], spp
    uniform in 1..spp_hi and maxBounces uniform in [depth_lo, depth_hi].
    Returns a list of dicts with the arrival time `at` in seconds."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    sides = np.arange(res_lo, res_hi + 1, 128)
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            break
        out.append({"at": t, "sceneIndex": int(rng.integers(0, n_scenes)),
                    "camera": orbit_camera(rng), "width": int(rng.choice(sides)),
                    "spp": int(rng.integers(1, spp_hi + 1)),
                    "maxBounces": int(rng.integers(depth_lo, depth_hi + 1)),
                    "seed": int(rng.integers(1, 1 << 30)), "russianRoulette": True, "nee": True})
        out[-1]["height"] = out[-1]["width"]
    return out


def uv_sphere(center, radius, bands=65, segments=128):
    """Tessellated sphere: 2*segments*(bands-1) triangles (16384 for the
    defaults), float32-rounded vertices, outward winding."""
    th = np.linspace(0, math.pi, bands + 1)
    ph = np.linspace(0, 2 * math.pi, segments + 1)[:-1]
    st, ct_ = np.sin(th), np.cos(th)
    ring = np.stack([st[:, None] * np.cos(ph)[None], np.broadcast_to(ct_[:, None], (bands + 1, segments)),
                     st[:, None] * np.sin(ph)[None]], -1)  # (bands+1, segments, 3), y up
    v = np.asarray(center, np.float64) + radius * ring
    j = np.arange(segments)
    j1 = (j + 1) % segments
    tris = []
    top, bot = v[0, 0], v[bands, 0]
    tris.append(np.stack([np.broadcast_to(top, (segments, 3)), v[1, j1], v[1, j]], 1))
    for b in range(1, bands - 1):
        tris.append(np.stack([v[b, j], v[b, j1], v[b + 1, j1]], 1))
        tris.append(np.stack([v[b, j], v[b + 1, j1], v[b + 1, j]], 1))
    tris.append(np.stack([np.broadcast_to(bot, (segments, 3)), v[bands - 1, j], v[bands - 1, j1]], 1))
    return _f32(np.concatenate(tris, 0))


def caustic_scene(n_spheres=256, sphere_bands=65, sphere_segments=128, plane_rows=128,
                  plane_cols=256, seed=3, light_area=0.01, light_radiance=500.0,
                  light_height=6.0):
    """Precision-form workload: n_spheres tessellated glass spheres (IOR 1.5,
    radii uniform in [0.2, 1.0]) resting on a checker-textured diffuse plane of
    2*plane_rows*plane_cols triangles, lit by one 2-triangle square light of
    the given area and radiance facing down; no sky.  The defaults give
    256 x 16384 + 65536 + 2 = 4,259,842 triangles."""
    rng = np.random.default_rng(seed)
    side = int(math.ceil(math.sqrt(n_spheres)))
    pitch = 2.5
    half = side * pitch / 2
    objects = []
    radii = rng.uniform(0.2, 1.0, n_spheres)
    for i in range(n_spheres):
        gx, gz = i % side, i // side
        r = float(radii[i])
        slack = pitch / 2 - r
        cx = -half + (gx + 0.5) * pitch + rng.uniform(-slack, slack)
        cz = -half + (gz + 0.5) * pitch + rng.uniform(-slack, slack)
        tri = uv_sphere((cx, r, cz), r, sphere_bands, sphere_segments)
        objects.append({"name": f"glass{i}", "material": {"type": "dielectric", "ior": 1.5},
                        "shape": {"kind": "mesh", "vertices": tri.reshape(-1, 3),
                                  "indices": np.arange(3 * len(tri))}})
    ext = half + 2.0
    xs = np.linspace(-ext, ext, plane_cols + 1)
    zs = np.linspace(-ext, ext, plane_rows + 1)
    X, Z = np.meshgrid(xs, zs)
    P = np.stack([X, np.zeros_like(X), Z], -1)
    a, b, c, d = P[:-1, :-1], P[:-1, 1:], P[1:, 1:], P[1:, :-1]
    plane = np.concatenate([np.stack([a, c, b], -2).reshape(-1, 3, 3),
                            np.stack([a, d, c], -2).reshape(-1, 3, 3)], 0)
    plane = _f32(plane)
    objects.append({"name": "plane", "material": {"type": "lambertian", "albedo": (0.8, 0.8, 0.8),
                                                  "checker": {"color2": (0.1, 0.1, 0.1),
                                                              "scale": 1.0}},
                    "shape": {"kind": "mesh", "vertices": plane.reshape(-1, 3),
                              "indices": np.arange(3 * len(plane))}})
    h = math.sqrt(light_area) / 2
    y = light_height
    q = _f32(np.array([[[-h, y, -h], [h, y, h], [h, y, -h]], [[-h, y, -h], [-h, y, h], [h, y, h]]]))
    objects.append({"name": "light", "material": {"type": "emissive", "color": (1.0, 1.0, 1.0),
                                                  "strength": light_radiance},
                    "shape": {"kind": "mesh", "vertices": q.reshape(-1, 3), "indices": np.arange(6)}})
    dist = 34.0
    camera = {"target": (0.0, 0.0, 0.0), "yaw": 0.35, "pitch": 0.6, "distance": dist,
              "fovDeg": 50.0, "lensRadius": 0.0, "focusDistance": dist}
    return {"name": "caustics", "objects": objects, "camera": camera, "sky": {"kind": "none"}}


def cornell_c1():
    """The small plain-method workload: buildCornellBox (scenes.ts:45-67) with
    its ceiling light at radiance 15 (white), the metal sphere at roughness 0.1
    and the back wall checker-modulated (color2 0.25, scale 8); rendered at
    256x256, 16 spp, maxBounces 8 over a 16-bucket SAH tree."""
    from .presets import cornell_box

    sc = cornell_box()
    objects = []
    for o in sc["objects"]:
        m = dict(o["material"])
        if o["name"] == "Ceiling light":
            m["color"], m["strength"] = (1, 1, 1), 15
        elif o["name"] == "Metal sphere":
            m["roughness"] = 0.1
        elif o["name"] == "Back wall":
            m["checker"] = {"color2": (0.25, 0.25, 0.25), "scale": 8}
        objects.append(dict(o, material=m))
    return dict(sc, objects=objects)


def bvh_test_rays(n, seed=99, lo=-14.0, hi=14.0, tlo=-8.0, thi=8.0):
    """The closest-hit ray set of bvh.test.ts:89-99 at any size: per ray, from
    Rng(seed), origin = range(lo, hi) x3, target = range(tlo, thi) x3,
    direction (target - origin) / Math.hypot(...) in binary64."""
    from .presets import hypot3, rng_stream

    d = rng_stream(seed, 6 * n).reshape(n, 6)
    o = lo + (hi - lo) * d[:, 0:3]
    t = tlo + (thi - tlo) * d[:, 3:6]
    v = t - o
    ln = hypot3(v[:, 0], v[:, 1], v[:, 2])
    return np.concatenate([o, v / ln[:, None]], 1)


def random_rays(rng_stream, n, lo=-14.0, hi=14.0, tlo=-8.0, thi=8.0):
    """Rays in the shape of bvh.test.ts:94-99 from a stream of uniforms in [0,1):
    origins in [lo,hi]^3 aimed at points of [tlo,thi]^3, unit binary64 directions."""
    d = np.asarray(rng_stream, np.float64).reshape(n, 6)
    o = lo + (hi - lo) * d[:, 0:3]
    t = tlo + (thi - tlo) * d[:, 3:6]
    v = t - o
    ln = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    return np.concatenate([o, v / ln[:, None]], 1)
