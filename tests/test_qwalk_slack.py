"""The 4-wide walk's plane test for binary64 rays, checked in exact arithmetic.

The binary64 kernels walk the quantized 4-wide nodes in FP32 (ct_qwalk.h):
the ray is rounded to float32 and each axis gets a slack D that covers the
distance between the binary64 ray and the float32 one at any parameter where
a hit can lie.  Closest hits stay the reference's only if every box a binary64
hit point lies in is entered with an entry no larger than the hit parameter.

Here the plane test is restated in numpy binary32 (an FMA is the binary64 sum
of an exact product, rounded to binary32) and run on adversarial cases: a
point on a face or corner of a small child box, a ray through it from up to
1e5 away (some with zero direction components), box origins up to 6.5e4 from
zero.  Whether the binary64 ray really passes through the box at t* is decided
with rationals; every such case must be entered with entry <= t* <= exit.  Two
planted mistakes -- no slack at all, and a slack left in world units instead
of units of t -- must each miss some cases, so the check has teeth.  (The
direction term of D is covered by the per-node slack wherever the hit lies
inside the axis's slab, as here; it matters on an axis whose zero direction
component the walk replaces by 1e-20.)
"""
from fractions import Fraction

import numpy as np

f32 = np.float32
SLAB_REL = f32(4.76837158203125e-7)
MARGIN = 1.0009765625


def fma32(a, b, c):
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def safe_inv(d):
    d = f32(d)
    if abs(d) < f32(1e-20):
        d = f32(-1e-20) if d < 0 else f32(1e-20)
    return f32(f32(1) / d)


def walk_ray(o, d, root_lo, root_hi, mistake=None):
    """qray64: float32 ray and per-axis slack."""
    of = [f32(x) for x in o]
    inv = [safe_inv(x) for x in d]
    c = [0.5 * (float(root_lo[a]) + float(root_hi[a])) for a in range(3)]
    h = [0.5 * (float(root_hi[a]) - float(root_lo[a])) for a in range(3)]
    M = sum(abs(o[a] - c[a]) for a in range(3)) + sum(abs(x) for x in h)
    nd = float(np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
    T = 2.0 * M / nd if nd > 0 else float("inf")
    D = []
    for a in range(3):
        deff = 1.0 / float(inv[a])
        delta = abs(o[a] - float(of[a])) + T * abs(d[a] - deff)
        if mistake == "no-slack":
            D.append(f32(0))
        elif mistake == "world-units":  # the distance itself, not converted to t
            D.append(f32(delta * MARGIN))
        else:
            D.append(f32(delta * abs(float(inv[a])) * MARGIN))
    return of, inv, D


def plane_test(of, inv, D, origin, scale, qlo, qhi):
    """qwide_step's test of one child: (entry, exit) in binary32."""
    tn, tf = [], []
    for a in range(3):
        B = f32(scale[a] * inv[a])
        A = f32(f32(origin[a] - of[a]) * inv[a])
        E = fma32(SLAB_REL, fma32(f32(255), abs(B), abs(A)), D[a])
        An, Af = f32(A - E), f32(A + E)
        qn, qf = (qhi[a], qlo[a]) if inv[a] < 0 else (qlo[a], qhi[a])
        tn.append(fma32(f32(qn), B, An))
        tf.append(fma32(f32(qf), B, Af))
    return max(max(tn), f32(0)), min(tf)


def cases(seed, n):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        e = int(rng.integers(-12, 1))
        scale = [f32(2.0 ** e)] * 3
        origin = [f32(rng.uniform(-65536, 65536) if rng.random() < 0.5 else rng.uniform(-20, 20))
                  for _ in range(3)]
        qlo = [int(rng.integers(0, 200)) for _ in range(3)]
        qhi = [q + int(rng.integers(0, 56)) for q in qlo]
        lo = [Fraction(float(origin[a])) + qlo[a] * Fraction(float(scale[a])) for a in range(3)]
        hi = [Fraction(float(origin[a])) + qhi[a] * Fraction(float(scale[a])) for a in range(3)]
        # a point on the box surface: each axis at the low face, the high face or inside
        x = []
        for a in range(3):
            r = rng.integers(0, 3)
            x.append(lo[a] if r == 0 else hi[a] if r == 1 else lo[a] + (hi[a] - lo[a]) / 3)
        d = rng.normal(size=3)
        z = rng.integers(0, 6)
        if z < 3:
            d[z] = 0.0
        d /= np.linalg.norm(d)
        t = float(10.0 ** rng.uniform(-2, 5))
        o = [float(float(x[a]) - t * d[a]) for a in range(3)]
        # keep the case only if the binary64 ray is inside the box at t exactly
        p = [Fraction(o[a]) + Fraction(t) * Fraction(float(d[a])) for a in range(3)]
        if not all(lo[a] <= p[a] <= hi[a] for a in range(3)):
            continue
        root_lo = [f32(float(lo[a]) - 1.0) for a in range(3)]
        root_hi = [f32(float(hi[a]) + 1.0) for a in range(3)]
        yield o, [float(v) for v in d], t, origin, scale, qlo, qhi, root_lo, root_hi


def run(mistake=None, seed=2024, n=3000):
    kept = missed = 0
    for o, d, t, origin, scale, qlo, qhi, rlo, rhi in cases(seed, n):
        of, inv, D = walk_ray(o, d, rlo, rhi, mistake)
        tn, tf = plane_test(of, inv, D, origin, scale, qlo, qhi)
        kept += 1
        if not (tn <= tf and Fraction(float(tn)) <= Fraction(t) <= Fraction(float(tf))):
            missed += 1
    return kept, missed


def test_binary64_rays_enter_every_box_they_pass():
    kept, missed = run()
    assert kept > 500
    assert missed == 0


def test_planted_mistakes_are_caught():
    for mistake in ("no-slack", "world-units"):
        kept, missed = run(mistake)
        assert kept > 500
        assert missed > 0, mistake
