#!/usr/bin/env python3
"""Throughput of the path-tracing hot path on one MI355X.

Workload (one step = one renderImage of the whole frame, all inputs resident
in HBM): synthetic soup of 1,000,000 equilateral FP32 triangles (centroids
uniform in [-10,10]^3, edge lognormal(ln 0.05, 0.8), random orientation;
materials 70% Lambertian / 20% metal / 10% dielectric over 1000 mesh objects)
plus 64 emissive triangles, 1920x1080, 4 spp, maxBounces 6, Russian roulette
and NEE + MIS on.  The BVH is built on the GPU (timed separately).

Prints one JSON line: Mrays/s (extension + shadow rays per second) as `value`,
spp*Mpix/s beside it, the roofline of the dominant kernel (closest-hit
extend) and a CPU baseline: the oracle's restatement of the reference tracer
on a bounded row sample of the same frame on the host cores.
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "Mrays/s & spp·Mpix/s at equal RMS error vs plain; p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0


def load_package():
    name = "caustic_tracer_amd"
    if name in sys.modules:
        return sys.modules[name]
    path = os.path.join(ROOT, "caustic-tracer_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location(
        name, path, submodule_search_locations=[os.path.dirname(path)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(ct, ps, frame, w, h, spp, depth, budget_s=12.0):
    """Oracle (CPU restatement of tracer.ts) on a band of rows of the same frame."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O  # the checker, timed as the CPU baseline

    import numpy as np

    view = O.Packed(ps.prims, ps.nodes, ps.nodeCount, ps.mats, ps.lights, ps.primObject,
                    ps.order, ps.sky)
    threads = max(1, min(16, os.cpu_count() or 1))
    acc = np.zeros(w * h * 3, np.float32)
    y0 = h // 2
    # probe: one pass over `threads` rows
    t = time.perf_counter()
    c = O.accumulate_pass(view, frame, w, h, 0, acc, max_bounces=depth, threads=threads,
                          rows=(y0, y0 + threads))
    probe = time.perf_counter() - t
    per_row_pass = probe / threads * threads  # wall seconds per `threads` rows
    rows = threads * max(1, int(budget_s / max(per_row_pass * spp, 1e-3)))
    rows = min(rows, h - y0)
    acc[:] = 0
    t = time.perf_counter()
    rays = 0
    for s in range(spp):
        c = O.accumulate_pass(view, frame, w, h, s, acc, max_bounces=depth, threads=threads,
                              rows=(y0, y0 + rows))
        rays += int(c[0]) + int(c[1])
    dt = time.perf_counter() - t
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"rows {y0}..{y0 + rows - 1} of the {w}x{h} frame, {spp} passes, "
                      f"{rays} rays in {dt:.1f} s ({rows * w * spp / dt / 1e6:.4f} spp*Mpix/s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=4)
    ap.add_argument("--depth", type=int, default=6)
    ap.add_argument("--walk", choices=["ordered", "stackless", "wide"], default="wide")
    ap.add_argument("--shadow-walk", choices=["ordered", "stackless", "wide"], default="wide")
    ap.add_argument("--stack", type=int, default=24, help="LDS stack entries of the ordered walks")
    ap.add_argument("--persist-blocks", type=int, default=2048,
                    help="workgroups of the persistent traversal launches (0: one lane per ray)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one process, one "
              "MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    ct = load_package()
    ct.init(0)
    w, h, spp, depth = args.width, args.height, args.spp, args.depth
    t = time.perf_counter()
    scene = ct.synth.soup_scene(args.tris)
    log(f"scene generated in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    ps = ct.packScene(scene)
    log(f"packed ({ps.primCount} prims, {ps.nodeCount} nodes) in {time.perf_counter() - t:.1f} s")

    # GPU BVH build time on device-resident PrimBounds
    kind, a, b, c, r, obj = ct.pack._expand(scene)
    mn, mx, cen = ct.primBounds(kind, a, b, c, r)
    d_b = [ct.DeviceBuffer.from_array(x) for x in (mn, mx, cen)]
    ct.buildBvh(*d_b)  # warm
    bt = []
    for _ in range(3):
        ct.sync()
        t = time.perf_counter()
        nodes, nc, order = ct.buildBvh(*d_b)
        bt.append((time.perf_counter() - t) * 1e3)
    build_ms = min(bt)
    assert nc == ps.nodeCount and np.array_equal(nodes.view(np.uint32), ps.nodes.view(np.uint32))
    log(f"GPU BVH build {build_ms:.1f} ms (best of 3, includes copy-back)")
    del d_b

    ds = ct.DeviceScene(ps)
    frame = ct.cameraFrame(scene["camera"], w / h)
    walks = {"ordered": ct.WALK_ORDERED, "stackless": ct.WALK_STACKLESS, "wide": ct.WALK_WIDE}
    walk = walks[args.walk]
    tr = ct.Tracer(walk=walk)
    tr.set_traversal(walks[args.shadow_walk], args.stack)
    tr.set_schedule(args.persist_blocks > 0, max(1, args.persist_blocks))
    out = ct.DeviceBuffer.empty((w * h * 3,), np.float32)
    opts = {"width": w, "height": h, "spp": spp, "seed": 1, "maxBounces": depth,
            "russianRoulette": True, "nee": True}
    for i in range(args.warmup):
        tr.render_image(ds, frame, opts, out)
        ct.sync()
        log(f"warmup {i} done")
    ct.sync()
    t = time.perf_counter()
    for i in range(args.steps):
        tr.render_image(ds, frame, opts, out)
    ct.sync()
    total = time.perf_counter() - t
    st = tr.stats()
    rays = st["extensionRays"] + st["shadowRays"]
    ms_step = total / args.steps * 1e3
    mrays = rays * args.steps / total / 1e6
    sppmpix = w * h * spp * args.steps / total / 1e6
    log(f"timed {args.steps} steps: {ms_step:.1f} ms/step, {mrays:.0f} Mrays/s, "
        f"{sppmpix:.1f} spp*Mpix/s, rays/step ext {st['extensionRays']} shadow {st['shadowRays']}")

    # per-launch device time of the dominant kernel (HIP events on the launch stream)
    tr.set_profiling(True)
    tr.render_image(ds, frame, opts, out)
    prof = tr.stats()
    tr.set_profiling(False)
    # algorithmic work of the same launches (counting build of the kernels, untimed)
    tr.set_counting(True)
    tr.render_image(ds, frame, opts, out)
    cnt = tr.stats()
    tr.set_counting(False)
    launches = max(1, prof["extendLaunches"])
    ext_bytes = (32 * cnt["nodeVisits"] + 48 * cnt["triTests"] + 32 * cnt["sphereTests"] +
                 (24 + 16) * cnt["extensionRays"])
    ext_s = prof["extendMs"] / 1e3
    achieved = ext_bytes / ext_s / 1e9 if ext_s > 0 else 0.0
    log(f"extend: {launches} launches, {prof['extendMs']:.2f} ms total, "
        f"{cnt['nodeVisits'] / max(1, cnt['extensionRays']):.1f} nodes/ray, "
        f"{cnt['triTests'] / max(1, cnt['extensionRays']):.1f} tris/ray, {achieved:.0f} GB/s, "
        f"{cnt['stackFallbacks']} stack fallbacks; shade {prof['shadeMs']:.2f} ms, "
        f"shadow {prof['shadowMs']:.2f} ms "
        f"({cnt['shadowNodeVisits'] / max(1, cnt['shadowRays']):.1f} nodes/ray, "
        f"{cnt['shadowStackFallbacks']} fallbacks)")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"k_extend<{args.walk},{args.stack}>",
            "bytes_per_launch": ext_bytes // launches,
            "avg_launch_ms": round(prof["extendMs"] / launches, 4)}
    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle on host cores)...")
        cpu = cpu_baseline(ct, ps, frame, w, h, spp, depth, args.cpu_budget)
        log(f"CPU baseline {cpu['value']:.3f} Mrays/s on {cpu['cores']} threads")
    res = {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"triangle soup {args.tris} tris + 64 emissive, {w}x{h}, "
                               f"{spp} spp, maxBounces {depth}, RR+NEE, walk {args.walk}, "
                               f"shadow walk {args.shadow_walk}, stack {args.stack}, "
                               f"persistent blocks {args.persist_blocks}",
                   "triangles": args.tris, "width": w, "height": h, "spp": spp,
                   "max_bounces": depth, "parallelism": "single GPU"},
        "spp_mpix_per_s": round(sppmpix, 3),
        "rays_per_step": rays, "bvh_build_ms": round(build_ms, 2),
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
