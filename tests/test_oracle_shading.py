"""materials.test.ts, furnace.test.ts and reference.test.ts:61-94 restated
against the oracle: closed forms and statistical moments."""
import math

import numpy as np
import pytest

import scenes


def _norm(v):
    v = np.asarray(v, float)
    return v / math.sqrt(float(v @ v))


def test_cosine_hemisphere_moments(O):  # materials.test.ts:16-64
    rng = O.Rng(42)
    n = _norm([0.3, 0.8, -0.5])
    N = 100_000
    u = rng.take(2 * N).reshape(N, 2)
    ds = np.array([O.cosine_hemisphere(n, a, b) for a, b in u])
    assert np.all(np.abs(np.linalg.norm(ds, axis=1) - 1) < 1e-6)
    c = ds @ n
    assert (c >= 0).all()
    assert abs(c.mean() - 2 / 3) / (2 / 3) < 0.02
    hist = np.histogram(np.minimum(c, 0.999999), bins=10, range=(0, 1))[0]
    for b in range(10):
        lo, hi = b / 10, (b + 1) / 10
        exp = (hi * hi - lo * lo) * N
        assert abs(hist[b] - exp) / exp < 0.08


def test_dielectric_closed_forms(O):  # materials.test.ts:66-111
    expected = ((1.5 - 1) / (1.5 + 1)) ** 2
    assert abs(O.schlick(1, 1, 1.5) - expected) < 1e-6
    assert abs(O.dielectric_reflectance(1, 1, 1.5) - expected) < 1e-6
    assert abs(O.schlick(0, 1, 1.5) - 1) < 1e-6
    crit = math.asin(1 / 1.5)
    th = crit + 0.05
    d = _norm([math.sin(th), -math.cos(th), 0])
    assert O.dielectric_reflectance(math.cos(th), 1.5, 1) == 1
    _, refl = O.dielectric(d, (0, 1, 0), 1.5, False, 0.999)
    assert refl
    # Snell at 30 degrees, entering IOR 1.5 (u = 1 forces refraction)
    th = math.radians(30)
    d = _norm([math.sin(th), -math.cos(th), 0])
    t, refl = O.dielectric(d, (0, 1, 0), 1.5, True, 1.0)
    assert not refl
    assert abs(math.acos(min(1, -t[1])) - math.asin(math.sin(th) / 1.5)) < 1e-4
    assert abs(np.linalg.norm(t) - 1) < 1e-6 and abs(t[2]) < 1e-9
    # Fresnel split
    rng = O.Rng(3)
    d = _norm([0.5, -1, 0.2])
    R = O.dielectric_reflectance(-d[1], 1, 1.5)
    us = rng.take(50_000)
    refl = sum(O.dielectric(d, (0, 1, 0), 1.5, True, u)[1] for u in us)
    assert abs(refl / 50_000 - R) < 0.01


def test_metal(O):  # materials.test.ts:113-139
    rng = O.Rng(1)
    d = _norm([0.4, -0.7, 0.2])
    n = _norm([0.1, 1, 0.05])
    u = rng.take(3)
    r = O.metal(d, n, 0, *u)
    exp = d - 2 * float(d @ n) * n
    assert np.all(np.abs(r - exp) < 1e-12)
    assert abs(float(r @ n) + float(d @ n)) < 1e-12
    rng = O.Rng(5)
    s = 0
    for _ in range(1000):
        s += O.metal(np.array([0, -1.0, 0]), np.array([0, 1.0, 0]), 0.3, *rng.take(3))[1]
    assert s / 1000 > 0.9


def test_sampling_helpers(O):  # materials.test.ts:141-160
    rng = O.Rng(11)
    up = 0
    for _ in range(20_000):
        v = O.unit_vector(rng.next(), rng.next())
        assert abs(np.linalg.norm(v) - 1) < 1e-9
        up += v[2] > 0
        r = math.sqrt(rng.next())
        th = 2 * math.pi * rng.next()
        assert (r * math.cos(th)) ** 2 + (r * math.sin(th)) ** 2 <= 1 + 1e-12
    assert abs(up / 20_000 - 0.5) < 0.02
    for a, b in [(0.3, 0.9), (2, 0.1), (1, 1)]:
        assert abs(O.power_heuristic(a, b) + O.power_heuristic(b, a) - 1) < 1e-12
    assert O.power_heuristic(0, 0) == 0


def test_checker_parity(O):
    # materials.ts:177-181: floor(p*s+0.5) summed, even -> primary
    assert O.checker_primary((0.1, 0.1, 0.1), 1)
    assert not O.checker_primary((0.6, 0.1, 0.1), 1)
    assert O.checker_primary((-0.6, 0.6, 0.1), 1)


@pytest.mark.slow
def test_furnace_naive(O):  # furnace.test.ts:28-43
    sc = scenes.furnace()
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 1.0)
    img = O.render_image(ps, fr, 8, 8, 4096, seed=3, max_bounces=64, russian_roulette=False,
                         nee=False)
    assert np.all(np.abs(img - 5) / 5 < 0.05)


@pytest.mark.slow
def test_furnace_nee_mis(O):  # furnace.test.ts:45-63
    sc = scenes.furnace()
    ps = O.pack_scene(sc)
    assert len(ps.lights) == 1
    fr = O.camera_frame(sc["camera"], 1.0)
    img = O.render_image(ps, fr, 4, 4, 4096, seed=9, max_bounces=64, russian_roulette=False,
                         nee=True)
    assert abs(img.mean() - 5) / 5 < 0.03
    assert np.all(np.abs(img - 5) / 5 < 0.08)


@pytest.mark.slow
def test_furnace_russian_roulette(O):  # furnace.test.ts:65-81
    sc = scenes.furnace()
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 1.0)
    img = O.render_image(ps, fr, 4, 4, 4096, seed=5, max_bounces=64, russian_roulette=True,
                         nee=False)
    assert abs(img.mean() - 5) / 5 < 0.03


def test_tonemap_closed_forms(O):  # reference.test.ts:61-94
    pa = pr = -1
    for x in np.arange(0, 20.0001, 0.1):
        a, r = O.aces(x), O.reinhard(x)
        assert a >= pa - 1e-9 and r >= pr - 1e-9 and a <= 1 and r < 1
        pa, pr = a, r
    assert O.aces(0) == 0 and O.reinhard(0) == 0 and abs(O.reinhard(1) - 0.5) < 1e-12
    assert O.linear_to_srgb(0) == 0 and O.linear_to_srgb(1) == 1
    assert abs(O.linear_to_srgb(0.5) - 0.7354) < 5e-4
    acc = np.array([2, 2, 2], np.float32)
    a = O.tonemap_buffer(acc, 1, 1, 4, 0, "reinhard")
    assert a[0] == round(O.linear_to_srgb(O.reinhard(0.5)) * 255) and a[3] == 255
    b = O.tonemap_buffer(acc, 1, 1, 4, 1, "reinhard")
    assert b[0] == round(O.linear_to_srgb(O.reinhard(1.0)) * 255)


def test_rng_known_answers(O):
    # rng.ts: pcgHash(0) from the LCG step state = 2891336453 (closed form by hand)
    state = 2891336453
    shifted = ((state >> ((state >> 28) + 4)) ^ state) & 0xFFFFFFFF
    word = (shifted * 277803737) & 0xFFFFFFFF
    assert O.pcg_hash(0) == ((word >> 22) ^ word)
    # xorshift32 stream from Rng(seed): state0 = pcgHash(seed)
    s = O.pcg_hash(77)
    out = []
    for _ in range(5):
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        out.append(s)
    assert list(O.rng_stream_u32(77, 5)) == out
    # mean of next() over a long stream
    assert abs(O.rng_stream(5, 200_000).mean() - 0.5) < 0.005
    # seedFor zero guard and mixing (rng.ts:18-23)
    a = O.pcg_hash(3)
    b = O.pcg_hash(9 ^ 0x9E3779B9)
    sd = O.pcg_hash((a ^ ((b * 0x85EBCA6B) & 0xFFFFFFFF)) & 0xFFFFFFFF)
    assert O.seed_for(3, 9) == (sd if sd else 0x1234567)


def test_render_image_sample_count_edges(O):
    """renderImage (tracer.ts:348-355) divides the sums by spp: one sample
    leaves the accumulatePass sum unchanged, and zero samples leave NaN
    everywhere (zero sums times 1/0), which the GPU path reproduces."""
    sc = scenes.cornell_box()
    ps = O.pack_scene(sc)
    w, h = 6, 4
    frame = O.camera_frame(sc["camera"], w / h)
    one = O.render_image(ps, frame, w, h, 1, seed=3, max_bounces=3)
    acc = np.zeros(w * h * 3, np.float32)
    O.accumulate_pass(ps, frame, w, h, 3 * 7919, acc, max_bounces=3)
    assert np.array_equal(one.view(np.uint32), acc.view(np.uint32))
    none = O.render_image(ps, frame, w, h, 0, seed=3, max_bounces=3)
    assert none.shape == (w * h * 3,) and np.isnan(none).all()


def test_render_pixels_equal_render_image(O):
    """renderImage's values at listed pixels (the CPU reference of the
    precision form) equal those of the whole image, bit for bit."""
    import numpy as np

    import scenes
    sc = scenes.cornell_box()
    ps = O.pack_scene(sc)
    fr = O.camera_frame(sc["camera"], 1.5)
    img = O.render_image(ps, fr, 24, 16, 5, seed=3).reshape(-1, 3)
    px = np.array([0, 5, 17, 100, 383, 200])
    got = O.render_pixels(ps, fr, 24, 16, px, 5, seed=3, threads=3)
    assert np.array_equal(got.view(np.uint32), img[px].view(np.uint32))
