"""The product's host packer (pack.ts restated with numpy) against the oracle's
restatement: primitive records, normals, areas, PrimBounds, materials, lights.
The BVH part of packScene runs on the GPU and is covered by the GPU tests."""
import numpy as np

import scenes


def _host_pack(ct, sc, O):
    k, a, b, c, r, obj = ct.pack._expand(sc)
    recs = ct.pack._prim_records(k, a, b, c, r, obj)
    mn, mx, cen = ct.primBounds(k, a, b, c, r)
    return recs, mn, mx, cen, obj


def test_records_and_bounds_match_oracle(ct, O):
    for sc in (scenes.random_scene(O.Rng(1234), 1000, 2000), scenes.cornell_box(),
               scenes.night_emissives(), scenes.glass_spheres(O.Rng(20240607)),
               ct.synth.soup_scene(3000, n_materials=7)):
        recs, mn, mx, cen, obj = _host_pack(ct, sc, O)
        ps = O.pack_scene(sc)
        kind, geom, oobj = O.expand_scene(sc)
        bmin, bmax, bcen = O.prim_bounds(kind, geom)
        assert np.array_equal(mn.reshape(-1), bmin)
        assert np.array_equal(mx.reshape(-1), bmax)
        assert np.array_equal(cen.reshape(-1), bcen)
        n = ps.primCount
        assert np.array_equal(recs[ps.order].reshape(-1).view(np.uint32),
                              ps.prims[: 16 * n].view(np.uint32))
        mats = np.zeros(len(sc["objects"]) * 16, np.float32)
        for i, ob in enumerate(sc["objects"]):
            ct.pack.write_material(mats, i, ob["material"])
        assert np.array_equal(mats.view(np.uint32), ps.mats.view(np.uint32))


def test_camera_frame_matches_oracle(ct, O):
    for sc in (scenes.cornell_box(), scenes.glass_spheres(O.Rng(20240607)),
               scenes.night_emissives()):
        for aspect in (1.0, 1.5, 16 / 9):
            a = ct.cameraFrame(sc["camera"], aspect)
            b = O.camera_frame(sc["camera"], aspect)
            assert np.array_equal(a, b)
