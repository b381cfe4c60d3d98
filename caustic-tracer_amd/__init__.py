"""caustic-tracer hot path on MI355X (gfx950).

The public names mirror the reference's barrel (src/tracer/index.ts:1-20) for
the path this package accelerates: packScene / buildBvh (pack.ts, bvh.ts),
intersectClosest / occluded (intersect.ts), accumulatePass / renderImage
(tracer.ts) and cameraFrame (camera.ts).  All tracing and the BVH build run in
HIP kernels through libcaustic_amd.so; importing works without a GPU, and any
call that needs the library fails loudly when it is missing.
"""
from . import display, presets, scene_io, serve, synth
from .presets import SCENE_PRESETS, Rng, presetById
from .presets import pcg_hash as pcgHash
from .presets import seed_for as seedFor
from .scene_io import SceneParseError, parseScene, parseSceneJSON, sceneToJSON
from .display import decodePPM, encodePPM, tonemapBuffer
from ._capi import CTError, DeviceBuffer, init, sync
from .pack import PackedScene, buildBvh, packScene, primBounds
from .tracer import (PATHS_AUTO, PATHS_PER_PATH, PATHS_WAVEFRONT, PRECISION_F32, PRECISION_F64,
                     T_MIN, WALK_BRUTE, WALK_ORDERED,
                     WALK_STACKLESS, WALK_WIDE, DeviceScene, Tracer, accumulatePass, cameraFrame,
                     closest_hit, closest_hit32, closest_hit32_counts, closest_hit_counts,
                     image_sum_f64, intersectClosest, occluded, occluded32, pick, primaryRay,
                     radiance_batch, radianceInto, renderImage)
from .serve import Server

__all__ = [
    "presets", "SCENE_PRESETS", "presetById", "scene_io", "SceneParseError", "parseScene",
    "parseSceneJSON", "sceneToJSON", "display", "tonemapBuffer", "encodePPM", "decodePPM", "serve", "Server", "synth", "CTError", "DeviceBuffer", "init", "sync", "PackedScene", "buildBvh", "packScene",
    "primBounds", "T_MIN", "WALK_ORDERED", "WALK_STACKLESS", "WALK_WIDE", "WALK_BRUTE",
    "PRECISION_F64", "PRECISION_F32", "PATHS_AUTO", "PATHS_WAVEFRONT", "PATHS_PER_PATH",
    "closest_hit32", "closest_hit32_counts", "occluded32",
    "image_sum_f64", "pick", "primaryRay", "DeviceScene", "Tracer",
    "accumulatePass", "cameraFrame", "closest_hit", "closest_hit_counts", "intersectClosest",
    "occluded", "radianceInto", "radiance_batch", "renderImage", "Rng", "pcgHash", "seedFor",
]
