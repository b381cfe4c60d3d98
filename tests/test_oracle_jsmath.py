"""Math.sin, Math.cos and Math.cbrt of the oracle against the JS engine's own
outputs (tests/golden/jsmath_vectors.json, written by
make_jsmath_vectors.js under node): bit for bit on the arguments the tracer
passes (2*pi*u and u, u = k/2^32) and on a spread of others (multiples of
pi/2 and their neighbours, kernel thresholds, tiny, subnormal and huge values).
The C library's functions, which the oracle used before, miss on a few percent
of the trigonometric vectors -- the check has teeth."""
import json
import math
import os
import struct

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _vectors():
    with open(os.path.join(HERE, "golden", "jsmath_vectors.json")) as f:
        return json.load(f)


def _d(h):
    return struct.unpack(">d", bytes.fromhex(h))[0]


def _bits(x):
    return struct.pack(">d", x).hex()


@pytest.mark.parametrize("name", ["sin", "cos", "cbrt"])
def test_engine_math_bit_exact(O, name):
    fn = {"sin": O.js_sin, "cos": O.js_cos, "cbrt": O.js_cbrt}[name]
    vec = _vectors()[name]
    bad = [(x, y, _bits(fn(_d(x)))) for x, y in vec if _bits(fn(_d(x))) != y]
    assert len(vec) > 3000 and bad == []


def test_c_library_differs_from_the_engine():
    v = _vectors()
    for name, fn in (("sin", math.sin), ("cos", math.cos)):
        miss = sum(_bits(fn(_d(x))) != y for x, y in v[name])
        assert miss > 20, name  # the vectors separate the two implementations


def test_hypot_restatements_bit_exact(ct):
    """Math.hypot(x, y, z) (the ray sets' normalisation, bvh.test.ts:97): the
    scalar restatements of the tests and the presets and the vector form that
    generates the 16M-ray set all equal the engine's results."""
    import numpy as np

    import scenes
    vec = _vectors()["hypot3"]
    xs = np.array([[_d(a), _d(b), _d(c)] for a, b, c, _ in vec])
    want = [r for _, _, _, r in vec]
    v = ct.presets.hypot3(xs[:, 0], xs[:, 1], xs[:, 2])
    assert [_bits(float(x)) for x in v] == want
    assert [_bits(scenes.js_hypot(*row)) for row in xs] == want
    assert [_bits(ct.presets._hypot(*row)) for row in xs] == want
