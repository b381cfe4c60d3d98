"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/caustic_amd.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "caustic_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ct_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("ct_scene_create", "ct_closest_hit", "ct_occluded", "ct_accumulate_pass",
                 "ct_render_image", "ct_bvh_build"):
        assert must in names


def test_library_exports_every_declared_symbol(ct):
    lib = ctypes.CDLL(ct._capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header(ct):
    assert set(declared_functions()) == set(ct._capi.SIGNATURES)


def test_errors_without_gpu_are_loud(ct):
    # the binding reports status codes as exceptions; a bad walk id is rejected
    # before any device work
    lib = ct._capi.lib()
    rc = lib.ct_tracer_set_walk(None, 0)
    assert rc != 0
    assert b"null" in lib.ct_last_error()


def _create(ct, nodes, primCount=1, matCount=1, lights=(), pobj=None):
    import numpy as np
    C = ctypes
    lib = ct._capi.lib()
    nodes = np.ascontiguousarray(nodes, np.float32)
    prims = np.zeros(16 * max(1, primCount), np.float32)
    mats = np.zeros(16 * max(1, matCount), np.float32)
    li = np.ascontiguousarray(lights, np.int32)
    po = np.zeros(max(1, primCount), np.uint32) if pobj is None else np.asarray(pobj, np.uint32)
    h = C.c_void_p()
    rc = lib.ct_scene_create(prims.ctypes.data_as(C.c_void_p), primCount,
                             nodes.ctypes.data_as(C.c_void_p), len(nodes) // 8,
                             mats.ctypes.data_as(C.c_void_p), matCount,
                             li.ctypes.data_as(C.c_void_p), len(li),
                             po.ctypes.data_as(C.c_void_p), None, C.byref(h))
    return rc, lib.ct_last_error().decode()


def test_scene_upload_rejects_malformed_arrays(ct):
    # validation runs on the host before any device allocation
    rc, msg = _create(ct, [0, 0, 0, 0, 1, 1, 1, -1])          # skip link not forward
    assert rc == -1 and "skip" in msg
    rc, msg = _create(ct, [0, 0, 0, 1, 1, 1, 1, 8 * 5 + 2])   # leaf beyond primCount
    assert rc == -1 and "beyond primCount" in msg
    rc, msg = _create(ct, [0, 0, 0, 1, 1, 1, 1, -1])          # interior without children
    assert rc == -1 and "children" in msg
    rc, msg = _create(ct, [0, 0, 0, 1, 1, 1, 1, 1], lights=[3])
    assert rc == -1 and "light" in msg
    rc, msg = _create(ct, [0, 0, 0, 1, 1, 1, 1, 1], pobj=[7])
    assert rc == -1 and "primObject" in msg
