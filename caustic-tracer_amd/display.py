"""Display side of a render (tonemap.ts, ppm.ts): the tone map runs on the
GPU (ct_tonemap), the P6 container is written and read on the host."""
import ctypes as C

import numpy as np

from . import _capi as A
from ._capi import DeviceBuffer, check, lib

MODES = {"aces": 0, "reinhard": 1}


def tonemapBuffer(accum, width, height, samples, exposureEv=0.0, mode="aces", out=None):
    """tonemapBuffer (tonemap.ts:41-62): float32 sum (w*h*3, host array or
    DeviceBuffer) -> RGBA8 (w*h*4 numpy uint8, or written into the DeviceBuffer
    `out`)."""
    A.init()
    d_acc = accum if isinstance(accum, DeviceBuffer) else DeviceBuffer.from_array(
        np.ascontiguousarray(accum, np.float32).reshape(-1))
    n = width * height
    d_out = out if isinstance(out, DeviceBuffer) else DeviceBuffer.empty((max(1, 4 * n),), np.uint8)
    check(lib().ct_tonemap(d_acc.ptr, width, height, float(samples), float(exposureEv),
                           MODES[mode], d_out.ptr, None), "ct_tonemap")
    if isinstance(out, DeviceBuffer):
        return out
    return d_out.numpy()[: 4 * n].copy()


def encodePPM(width, height, rgba):
    """encodePPM (ppm.ts:4-16): binary P6 from RGBA8, row 0 at the top."""
    px = np.asarray(rgba, np.uint8).reshape(-1, 4)[: width * height, :3]
    return f"P6\n{width} {height}\n255\n".encode() + px.tobytes()


def decodePPM(data):
    """decodePPM (ppm.ts:25-46): (width, height, rgb bytes as (h, w, 3) uint8)."""
    toks, i = [], 0
    while len(toks) < 4 and i < len(data):
        while i < len(data) and data[i] in b" \n\r\t":
            i += 1
        if i < len(data) and data[i] == 0x23:  # comment to end of line
            while i < len(data) and data[i] != 0x0A:
                i += 1
            continue
        j = i
        while j < len(data) and data[j] not in b" \n\r\t":
            j += 1
        toks.append(data[i:j].decode())
        i = j
    if not toks or toks[0] != "P6":
        raise ValueError("not a binary PPM")
    if toks[3] != "255":
        raise ValueError("only maxval 255 supported")
    w, h = int(toks[1]), int(toks[2])
    i += 1  # single whitespace after maxval
    rgb = np.frombuffer(data[i:i + w * h * 3], np.uint8)
    if rgb.size != w * h * 3:
        raise ValueError("truncated PPM data")
    return w, h, rgb.reshape(h, w, 3)
