"""The five preset scenes of scenes.ts (scenes.ts:45-237) and the torus-knot
mesh of mesh.ts:46-80, built as this package's scene dicts.  Random layouts
draw from a host restatement of the reference RNG (rng.ts:10-64) in the
generators' order, so the packed arrays equal the reference's."""
import math

import numpy as np

# ---------------------------------------------------------------------------
# rng.ts (host side; u32 arithmetic)

M32 = 0xFFFFFFFF


def pcg_hash(x):
    """pcgHash (rng.ts:10-15)."""
    state = ((x & M32) * 747796405 + 2891336453) & M32
    shifted = ((state >> ((state >> 28) + 4)) ^ state) & M32
    word = (shifted * 277803737) & M32
    return ((word >> 22) ^ word) & M32


def seed_for(pixel, frame):
    """seedFor (rng.ts:18-23)."""
    a = pcg_hash(pixel & M32)
    b = pcg_hash((frame & M32) ^ 0x9E3779B9)
    s = pcg_hash(a ^ ((b * 0x85EBCA6B) & M32))
    return 0x1234567 if s == 0 else s


class Rng:
    """Rng (rng.ts:28-64): PCG-seeded xorshift32."""

    def __init__(self, seed):
        s = pcg_hash(seed & M32)
        self.state = 0x9E3779B9 if s == 0 else s

    @staticmethod
    def fromPixel(pixel, frame):
        """Rng.fromPixel (rng.ts:33-37): the state is seedFor(pixel, frame)."""
        r = Rng(0)
        r.state = seed_for(pixel, frame)
        return r

    def next_u32(self):
        x = self.state
        x ^= (x << 13) & M32
        x ^= x >> 17
        x ^= (x << 5) & M32
        self.state = x & M32
        return self.state

    def next(self):
        return self.next_u32() / 4294967296.0

    def range(self, lo, hi):
        return lo + (hi - lo) * self.next()

    def int(self, n):
        """rng.ts:62-64: an integer in [0, n)."""
        return min(n - 1, int(self.next() * n // 1))

    nextU32 = next_u32


def _xorshift_step(x):
    x ^= (x << 13) & M32
    x ^= x >> 17
    x ^= (x << 5) & M32
    return x & M32


def _apply_map(cols, v):
    """A GF(2)-linear map of 32-bit words (cols[j] = image of bit j) applied to v."""
    r = 0
    j = 0
    while v:
        if v & 1:
            r ^= cols[j]
        v >>= 1
        j += 1
    return r


def _jump_tables(cols):
    """Byte tables of a linear map: map(v) = t0[v & 255] ^ t1[..] ^ t2[..] ^ t3[v >> 24]."""
    t = np.zeros((4, 256), np.uint32)
    for k in range(4):
        for b in range(256):
            t[k, b] = _apply_map(cols, b << (8 * k))
    return t


def rng_u32_stream(seed, n, block=4096):
    """The first n outputs of Rng(seed).nextU32() (rng.ts:28-50) as uint32.

    xorshift32 is linear over GF(2): with T the one-step map, output k + m is
    T^m of output k.  The first `block` outputs are stepped one by one, then the
    filled prefix is doubled by applying T^filled to it (byte lookup tables),
    so n outputs cost a few table passes over the array."""
    out = np.empty(max(n, 0), np.uint32)
    if n <= 0:
        return out
    x = Rng(seed).state
    first = min(n, block)
    for i in range(first):
        x = _xorshift_step(x)
        out[i] = x
    filled = first
    cols = [_xorshift_step(1 << j) for j in range(32)]  # T
    power = [c for c in cols]
    # T^block by repeated squaring
    acc = [1 << j for j in range(32)]
    e, base = block, power
    while e:
        if e & 1:
            acc = [_apply_map(base, c) for c in acc]
        base = [_apply_map(base, c) for c in base]
        e >>= 1
    jump = acc  # T^filled while filled == block
    while filled < n:
        t = _jump_tables(jump)
        take = min(filled, n - filled)
        src = out[:take]
        out[filled:filled + take] = (t[0][src & 0xFF] ^ t[1][(src >> 8) & 0xFF] ^
                                     t[2][(src >> 16) & 0xFF] ^ t[3][src >> 24])
        filled += take
        jump = [_apply_map(jump, c) for c in jump]  # T^(2 filled)
    return out


def rng_stream(seed, n):
    """The first n outputs of Rng(seed).next() (rng.ts:52-54), binary64."""
    return rng_u32_stream(seed, n).astype(np.float64) / 4294967296.0


def hypot3(x, y, z):
    """Math.hypot(x, y, z) elementwise over arrays, in the engine's order (the
    vector form of _hypot)."""
    a = [np.abs(np.asarray(v, np.float64)) for v in (x, y, z)]
    m = np.maximum(np.maximum(a[0], a[1]), a[2])
    safe = np.where(m > 0, m, 1.0)
    s = np.zeros_like(m)
    comp = np.zeros_like(m)
    with np.errstate(invalid="ignore"):  # inf / inf rows are replaced below
        for v in a:
            q = v / safe
            summand = q * q - comp
            prelim = s + summand
            comp = (prelim - s) - summand
            s = prelim
        r = np.sqrt(s) * m
    r = np.where(m == 0, 0.0, r)
    return np.where(np.isinf(m), np.inf, r)


def _hypot(*xs):
    """Math.hypot in V8's order: scaled by the largest magnitude, Kahan-summed."""
    a = [abs(float(x)) for x in xs]
    m = max(a) if a else 0.0
    if m == math.inf:
        return math.inf
    if m == 0:
        return 0.0
    s = comp = 0.0
    for x in a:
        n = x / m
        summand = n * n - comp
        prelim = s + summand
        comp = (prelim - s) - summand
        s = prelim
    return math.sqrt(s) * m


# ---------------------------------------------------------------------------
# scene.ts helpers

def _obj(name, shape, material):
    return {"name": name, "shape": shape, "material": material}


def _sphere(c, r):
    return {"kind": "sphere", "center": tuple(c), "radius": r}


def _quad(o, u, v):
    return {"kind": "quad", "origin": tuple(o), "u": tuple(u), "v": tuple(v)}


def _lambert(albedo, checker=None):
    m = {"type": "lambertian", "albedo": tuple(albedo)}
    if checker is not None:
        m["checker"] = checker
    return m


def _metal(albedo, roughness):
    return {"type": "metal", "albedo": tuple(albedo), "roughness": roughness}


def _glass(ior=1.5):
    return {"type": "dielectric", "ior": ior}


def _light(color, strength):
    return {"type": "emissive", "color": tuple(color), "strength": strength}


def camera_at(eye, target, fov_deg, lens_radius, focus_distance=None):
    """cameraAt (scenes.ts:26-41)."""
    dx, dy, dz = eye[0] - target[0], eye[1] - target[1], eye[2] - target[2]
    distance = math.sqrt(dx * dx + dy * dy + dz * dz)
    return {"target": tuple(target), "yaw": math.atan2(dx, dz), "pitch": math.asin(dy / distance),
            "distance": distance, "fovDeg": fov_deg, "lensRadius": lens_radius,
            "focusDistance": distance if focus_distance is None else focus_distance}


# ---------------------------------------------------------------------------
# mesh.ts

def _norm(v):
    ln = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    if ln > 0:
        s = 1 / ln
        return (v[0] * s, v[1] * s, v[2] * s)
    return (0.0, 0.0, 0.0)


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _knot_point(t, p, q, R, r):
    cq, sq = math.cos(q * t), math.sin(q * t)
    cp, sp = math.cos(p * t), math.sin(p * t)
    ring = R + r * cq
    return (ring * cp, r * sq, ring * sp), (R * cp, 0.0, R * sp)


def torus_knot(p, q, radius, ring_radius, tube, tubular_segments, radial_segments,
               center=(0.0, 0.0, 0.0)):
    """torusKnot (mesh.ts:46-80): a tube swept along a (p, q) torus knot, framed
    by the tangent and the torus-surface normal.  Returns (vertices (n,3), indices)."""
    verts, idx = [], []
    dt = 2 * math.pi / tubular_segments
    for i in range(tubular_segments):
        t = i * dt
        here, core = _knot_point(t, p, q, radius, ring_radius)
        nxt, _ = _knot_point(t + dt * 0.01, p, q, radius, ring_radius)
        tangent = _norm((nxt[0] - here[0], nxt[1] - here[1], nxt[2] - here[2]))
        normal = _norm((here[0] - core[0], here[1] - core[1], here[2] - core[2]))
        binormal = _norm(_cross(tangent, normal))
        normal = _cross(binormal, tangent)
        for j in range(radial_segments):
            a = (j / radial_segments) * 2 * math.pi
            ca, sa = math.cos(a) * tube, math.sin(a) * tube
            off = (normal[0] * ca + binormal[0] * sa, normal[1] * ca + binormal[1] * sa,
                   normal[2] * ca + binormal[2] * sa)
            verts.append(((here[0] + off[0]) + center[0], (here[1] + off[1]) + center[1],
                          (here[2] + off[2]) + center[2]))
    for i in range(tubular_segments):
        i1 = (i + 1) % tubular_segments
        for j in range(radial_segments):
            j1 = (j + 1) % radial_segments
            a, b = i * radial_segments + j, i1 * radial_segments + j
            c, d = i1 * radial_segments + j1, i * radial_segments + j1
            idx += [a, b, c, a, c, d]
    return np.array(verts, np.float64), np.array(idx, np.int64)


# ---------------------------------------------------------------------------
# scenes.ts

def cornell_box():
    """buildCornellBox (scenes.ts:45-67)."""
    white, red = _lambert((0.73, 0.73, 0.73)), _lambert((0.65, 0.05, 0.05))
    green = _lambert((0.12, 0.45, 0.15))
    W = 0.75
    objects = [
        _obj("Floor", _quad((-W, 0, -0.5), (2 * W, 0, 0), (0, 0, 1)), white),
        _obj("Ceiling", _quad((-W, 1, -0.5), (2 * W, 0, 0), (0, 0, 1)), white),
        _obj("Back wall", _quad((-W, 0, -0.5), (2 * W, 0, 0), (0, 1, 0)), white),
        _obj("Red wall", _quad((-W, 0, -0.5), (0, 0, 1), (0, 1, 0)), red),
        _obj("Green wall", _quad((W, 0, -0.5), (0, 0, 1), (0, 1, 0)), green),
        _obj("Ceiling light", _quad((-0.2, 0.998, -0.15), (0.4, 0, 0), (0, 0, 0.3)),
             _light((1, 0.86, 0.7), 12)),
        _obj("Metal sphere", _sphere((-0.33, 0.24, -0.12), 0.24), _metal((0.9, 0.9, 0.92), 0.02)),
        _obj("Glass sphere", _sphere((0.3, 0.2, 0.17), 0.2), _glass(1.5)),
    ]
    return {"name": "Cornell Box", "objects": objects,
            "camera": camera_at((0, 0.5, 1.6), (0, 0.5, 0), 32, 0, 1.6), "sky": {"kind": "none"}}


def glass_spheres():
    """buildGlassSpheres (scenes.ts:72-113): sixty seeded small spheres."""
    rng = Rng(20240607)
    objects = [
        _obj("Checker ground", _quad((-40, 0, -40), (80, 0, 0), (0, 0, 80)),
             _lambert((0.82, 0.8, 0.76), {"color2": (0.18, 0.17, 0.16), "scale": 1})),
        _obj("Big glass", _sphere((0, 1, 0), 1), _glass(1.5)),
        _obj("Big diffuse", _sphere((-4, 1, 0), 1), _lambert((0.4, 0.2, 0.1))),
        _obj("Big metal", _sphere((4, 1, 0), 1), _metal((0.7, 0.6, 0.5), 0)),
    ]
    bigs = [(0, 1, 0), (-4, 1, 0), (4, 1, 0)]
    n = 0
    for a in range(-6, 6):
        if n >= 60:
            break
        for b in range(-4, 5):
            if n >= 60:
                break
            c = (a * 1.15 + 0.7 * rng.next(), 0.2, b * 1.15 + 0.7 * rng.next())
            if any(_hypot(g[0] - c[0], g[2] - c[2]) < 1.4 for g in bigs):
                continue
            u = rng.next()
            if u < 0.62:
                m = _lambert((rng.next() * rng.next(), rng.next() * rng.next(),
                              rng.next() * rng.next()))
            elif u < 0.85:
                m = _metal((0.5 + 0.5 * rng.next(), 0.5 + 0.5 * rng.next(),
                            0.5 + 0.5 * rng.next()), 0.3 * rng.next())
            else:
                m = _glass(1.5)
            objects.append(_obj(f"Sphere {n + 1}", _sphere(c, 0.2), m))
            n += 1
    return {"name": "Glass Spheres on Checkerboard", "objects": objects,
            "camera": camera_at((12, 2.2, 3.2), (0, 0.55, 0), 22, 0.06, 10.5),
            "sky": {"kind": "gradient", "zenith": (0.5, 0.7, 1), "horizon": (1, 1, 1),
                    "strength": 1}}


def night_emissives():
    """buildNightEmissives (scenes.ts:118-142): lit only by glowing spheres."""
    glows = [((-1.6, 0.35, -0.6), 0.35, (1, 0.35, 0.1), 6), ((0.2, 0.25, 0.9), 0.25, (0.2, 0.55, 1), 7),
             ((1.7, 0.45, -1.1), 0.45, (0.6, 1, 0.3), 4), ((-0.4, 0.15, -1.9), 0.15, (1, 0.2, 0.6), 10),
             ((1.1, 0.18, 1.6), 0.18, (1, 0.85, 0.4), 8), ((-2.6, 0.2, 1.4), 0.2, (0.5, 0.3, 1), 7)]
    objects = [
        _obj("Dark ground", _quad((-30, 0, -30), (60, 0, 0), (0, 0, 60)), _lambert((0.09, 0.085, 0.08))),
        _obj("Chrome sphere", _sphere((0.4, 0.55, -0.7), 0.55), _metal((0.95, 0.95, 0.95), 0.0)),
        _obj("Brushed sphere", _sphere((-1.0, 0.4, 1.1), 0.4), _metal((0.9, 0.7, 0.5), 0.25)),
        _obj("Glass sphere", _sphere((1.9, 0.4, 0.6), 0.4), _glass(1.5)),
        _obj("Matte sphere", _sphere((-2.2, 0.5, -1.5), 0.5), _lambert((0.7, 0.7, 0.72))),
    ]
    for i, (p, r, c, s) in enumerate(glows):
        objects.append(_obj(f"Glow {i + 1}", _sphere(p, r), _light(c, s)))
    return {"name": "Night Emissives", "objects": objects,
            "camera": camera_at((4.2, 1.6, 4.6), (0, 0.35, 0), 30, 0.02, 6.2), "sky": {"kind": "none"}}


def depth_of_field():
    """buildDepthOfField (scenes.ts:147-180): a receding row under a wide aperture."""
    objects = [_obj("Checker ground", _quad((-30, 0, -30), (60, 0, 0), (0, 0, 60)),
                    _lambert((0.75, 0.74, 0.72), {"color2": (0.25, 0.24, 0.23), "scale": 2}))]
    palette = [_lambert((0.85, 0.3, 0.2)), _metal((0.9, 0.85, 0.7), 0.05), _glass(1.5),
               _lambert((0.2, 0.45, 0.85)), _metal((0.8, 0.8, 0.85), 0.2), _lambert((0.95, 0.8, 0.25)),
               _glass(1.5), _lambert((0.3, 0.7, 0.4)), _metal((0.95, 0.65, 0.5), 0),
               _lambert((0.9, 0.9, 0.9))]
    for i in range(10):
        objects.append(_obj(f"Sphere {i + 1}", _sphere((i * 0.42 - 1.0, 0.45, -i * 1.1), 0.45),
                            palette[i]))
    eye = (-1.6, 1.25, 3.6)
    focus_on = (-1.0 + 2 * 0.42, 0.45, -2.2)
    fd = _hypot(eye[0] - focus_on[0], eye[1] - focus_on[1], eye[2] - focus_on[2])
    return {"name": "Depth of Field Study", "objects": objects,
            "camera": camera_at(eye, (0.2, 0.45, -1.8), 30, 0.12, fd),
            "sky": {"kind": "gradient", "zenith": (0.45, 0.6, 0.9), "horizon": (0.95, 0.9, 0.85),
                    "strength": 0.9}}


def mirror_room():
    """buildMirrorRoom (scenes.ts:185-213): a 2000-triangle torus knot among mirrors."""
    wall = _metal((0.88, 0.86, 0.82), 0.015)
    verts, idx = torus_knot(2, 3, 0.55, 0.22, 0.11, 100, 10, (0, 1.0, 0))
    objects = [
        _obj("Floor", _quad((-2, 0, -2), (4, 0, 0), (0, 0, 4)), _lambert((0.16, 0.15, 0.14))),
        _obj("Ceiling", _quad((-2, 2.2, -2), (4, 0, 0), (0, 0, 4)), wall),
        _obj("Back wall", _quad((-2, 0, -2), (4, 0, 0), (0, 2.2, 0)), wall),
        _obj("Front wall", _quad((-2, 0, 2), (4, 0, 0), (0, 2.2, 0)), wall),
        _obj("Left wall", _quad((-2, 0, -2), (0, 0, 4), (0, 2.2, 0)), wall),
        _obj("Right wall", _quad((2, 0, -2), (0, 0, 4), (0, 2.2, 0)), wall),
        _obj("Light strip A", _quad((-1.2, 2.19, -0.15), (2.4, 0, 0), (0, 0, 0.3)), _light((1, 0.92, 0.8), 9)),
        _obj("Light strip B", _quad((-0.15, 2.19, -1.6), (0.3, 0, 0), (0, 0, 3.2)), _light((0.8, 0.9, 1), 5)),
        _obj("Torus knot", {"kind": "mesh", "vertices": verts, "indices": idx}, _lambert((0.95, 0.42, 0.12))),
    ]
    return {"name": "Mirror Room", "objects": objects,
            "camera": camera_at((1.15, 1.2, 1.75), (0, 0.95, 0), 40, 0.01, 2.4), "sky": {"kind": "none"}}


_counter = [0]


def _base36(n):
    digits = "0123456789abcdefghijklmnopqrstuvwxyz"
    out = ""
    while True:
        n, r = divmod(n, 36)
        out = digits[r] + out
        if n == 0:
            return out


def _with_ids(scene):
    """Object ids as scenes.ts:10-14 numbers them: p<n in base 36>, one counter
    over the presets in definition order."""
    for o in scene["objects"]:
        _counter[0] += 1
        o["id"] = f"p{_base36(_counter[0])}"
    return scene


CORNELL_BOX = _with_ids(cornell_box())
GLASS_SPHERES = _with_ids(glass_spheres())
NIGHT_EMISSIVES = _with_ids(night_emissives())
DEPTH_OF_FIELD = _with_ids(depth_of_field())
MIRROR_ROOM = _with_ids(mirror_room())

SCENE_PRESETS = [
    {"id": "cornell", "label": "Cornell Box", "scene": CORNELL_BOX},
    {"id": "glass", "label": "Glass Spheres", "scene": GLASS_SPHERES},
    {"id": "night", "label": "Night Emissives", "scene": NIGHT_EMISSIVES},
    {"id": "dof", "label": "Depth of Field", "scene": DEPTH_OF_FIELD},
    {"id": "mirror", "label": "Mirror Room", "scene": MIRROR_ROOM},
]


def presetById(pid):
    """presetById (scenes.ts:235-237)."""
    for p in SCENE_PRESETS:
        if p["id"] == pid:
            return p
    return None
