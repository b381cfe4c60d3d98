"""ct_bvh_build_noalloc: the BVH build in a static device buffer, through the C ABI.

Its output must equal ct_bvh_build_ex's bit for bit (nodes, order, exact leaf
words) on two calls in a row, the second with a smaller input (the buffer keeps
the first call's contents), and a count beyond the buffer is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _soup_bounds(O, ct, n, seed):
    tris = ct.synth.equilateral_soup(np.random.default_rng(seed), n)
    bmin, bmax, bcen = O.prim_bounds(np.ones(n, np.int32), tris.reshape(n, 9))
    return [ct.DeviceBuffer.from_array(np.ascontiguousarray(x).reshape(-1, 3))
            for x in (bmin, bmax, bcen)]


def _build(ct, fn, dev, n, bins):
    lib = ct._capi.lib()
    nodes = ct.DeviceBuffer.empty((max(1, 2 * n - 1) * 8,), np.float32)
    order = ct.DeviceBuffer.empty((max(1, n),), np.uint32)
    words = ct.DeviceBuffer.empty((max(1, 2 * n - 1),), np.uint32)
    nc = C.c_uint32(0)
    rc = getattr(lib, fn)(dev[0].ptr, dev[1].ptr, dev[2].ptr, n, bins, nodes.ptr, C.byref(nc),
                          order.ptr, words.ptr, None)
    assert rc == 0, ct._capi.lib().ct_last_error()
    k = nc.value
    return k, nodes.numpy()[: 8 * k].copy(), order.numpy()[:n].copy(), words.numpy()[:k].copy()


@pytest.mark.parametrize("bins", [12, 16])
def test_noalloc_build_equals_allocating_build(ct, O, bins):
    print(f"no-alloc BVH build, bins {bins}: 60000 then 3000 soup triangles", flush=True)
    for n, seed in ((60000, 5), (3000, 6)):  # the second call is the smaller one
        dev = _soup_bounds(O, ct, n, seed)
        a = _build(ct, "ct_bvh_build_ex", dev, n, bins)
        b = _build(ct, "ct_bvh_build_noalloc", dev, n, bins)
        assert a[0] == b[0]
        assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
        assert np.array_equal(a[2], b[2])
        assert np.array_equal(a[3], b[3])


def test_noalloc_build_refuses_counts_beyond_its_buffer(ct, O):
    lib = ct._capi.lib()
    dev = _soup_bounds(O, ct, 16, 7)  # the call fails before it reads them
    nodes = ct.DeviceBuffer.empty((8,), np.float32)
    order = ct.DeviceBuffer.empty((1,), np.uint32)
    nc = C.c_uint32(0)
    rc = lib.ct_bvh_build_noalloc(dev[0].ptr, dev[1].ptr, dev[2].ptr, 200000, 12, nodes.ptr,
                                  C.byref(nc), order.ptr, None, None)
    assert rc == -4  # CT_ELIMIT
    assert b"static buffer" in lib.ct_last_error()
