"""The oracle's end-to-end Cornell render equals the reference's own golden file.

tests/golden/cornell_96x64_16spp.ppm is out/cornell.ppm from the reference tree
(written by scripts/render-cpu.ts:13-28 via renderCornellReference, reference.ts:20-43:
96x64, 16 spp, maxBounces 8, RR and NEE on, seed 1, ACES, EV 0).  Matching it byte
for byte pins the RNG stream, packing, BVH build, camera, ordered walk, NEE + MIS,
materials, Russian roulette and the tone map at once.
"""
import os

import numpy as np

import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cornell_96x64_16spp.ppm")


def read_ppm(path):
    """P6 reader (header: magic, width, height, maxval; one whitespace byte)."""
    data = open(path, "rb").read()
    toks, i = [], 0
    while len(toks) < 4:
        while data[i:i + 1].isspace():
            i += 1
        if data[i:i + 1] == b"#":
            while data[i:i + 1] != b"\n":
                i += 1
            continue
        j = i
        while not data[j:j + 1].isspace():
            j += 1
        toks.append(data[i:j])
        i = j
    assert toks[0] == b"P6" and toks[3] == b"255"
    w, h = int(toks[1]), int(toks[2])
    rgb = np.frombuffer(data[i + 1:i + 1 + w * h * 3], np.uint8)
    assert rgb.size == w * h * 3
    return w, h, rgb.reshape(h, w, 3)


def test_golden_file_shape():
    w, h, rgb = read_ppm(GOLDEN)
    assert (w, h) == (96, 64)
    # reference.test.ts:18-48 qualitative checks hold on the golden data itself
    assert rgb[24:40, 1:7, 0].astype(int).sum() > 1.5 * rgb[24:40, 1:7, 1].astype(int).sum()
    assert rgb[24:40, 89:95, 1].astype(int).sum() > 1.5 * rgb[24:40, 89:95, 0].astype(int).sum()


def test_oracle_matches_golden_ppm_bytewise(O):
    sc = scenes.cornell_box()
    ps = O.pack_scene(sc)
    frame = O.camera_frame(sc["camera"], 96 / 64)
    lin = O.render_image(ps, frame, 96, 64, 16, seed=1, max_bounces=8)
    rgba = O.tonemap_buffer(lin, 96, 64, 1, 0.0, "aces").reshape(64, 96, 4)
    _, _, gold = read_ppm(GOLDEN)
    diff = np.abs(rgba[:, :, :3].astype(int) - gold.astype(int))
    assert diff.max() == 0, f"{(diff > 0).sum()} bytes differ, max {diff.max()}"
    # reference.test.ts:15-16 centre brighter than the bottom-left corner
    lum = lambda p: 0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2]
    L = lin.reshape(64, 96, 3)
    assert lum(L[32, 48]) > lum(L[63, 0])
    assert np.isfinite(lin).all() and lin.sum() > 0
