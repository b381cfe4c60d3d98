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
on a bounded row sample of the same frame on the host cores.  After the
frame, the line's "forms" object summarizes shorter runs of the other
configurations (--forms): the 256x256 Cornell variant with its whole-frame
parity and CPU rate, the precision, batching and serving forms below.

--form batch: 64 soup scenes of lognormal size (1k-500k triangles, median
20k), resolutions from {128,256,512,768}^2, 1-32 spp, maxBounces 8, traced
together after the launch settings are autotuned, replayed as a hipGraph;
beside it the same jobs rendered one after another with renderImage.
--form precision: 256 tessellated glass spheres (16384 triangles each) on a
65536-triangle checker plane under a 0.01-area light of radiance 500
(4,259,842 triangles), 1024x1024, 64 spp, maxBounces 16, rendered by the
precision form (quantized nodes, FP32, compensated sums) and by the binary64
form; each image's RMS error against a binary64 render at --ref-spp.
--form c1: the 256x256, 16 spp, maxBounces 8 Cornell variant alone (the
small plain workload; rounds over such scenes run the per-path launch).
--form serve: 128 resident scenes (1k-1M triangles) and a Poisson request
stream (random camera, 128^2-1024^2, 1-64 spp, maxBounces 4-12) through the
serving loop; p50/p99 time-to-complete and aggregate throughput, beside the
same requests rendered one at a time.
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
# VALU lane-op peak: 256 CUs x 64 lanes x 2.4 GHz, one non-packed lane-op per lane per cycle
VALU_PEAK_TLOPS = 39.3


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cpus():
    """Cores this process may use: the affinity mask, capped by a cgroup CPU
    quota when one is set (a container's share of a larger host)."""
    n = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = n
    quota = None
    for path, parse in (("/sys/fs/cgroup/cpu.max", lambda t: t.split()),
                        ("/sys/fs/cgroup/cpu/cpu.cfs_quota_us", None)):
        try:
            with open(path) as f:
                text = f.read().strip()
            if parse:
                q, per = parse(text)
                if q != "max":
                    quota = int(q) / int(per)
            else:
                q = int(text)
                with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
                    per = int(f.read().strip())
                if q > 0:
                    quota = q / per
            break
        except (OSError, ValueError):
            continue
    allowed = aff if quota is None else max(1, min(aff, int(quota)))
    return {"cpu_count": n, "affinity": aff, "cgroup_quota": quota, "allowed": allowed}


def _band_rates(O, view, frame, w, h, spp, depth, seed, threads, rows, acc, trials=3):
    """Trials of renderImage's passes (frame index s + 7919*seed) over a band."""
    rates, rays = [], 0
    for _ in range(trials):
        acc[rows[0] * w * 3:rows[1] * w * 3] = 0
        t = time.perf_counter()
        rays = 0
        for s in range(spp):
            c = O.accumulate_pass(view, frame, w, h, s + 7919 * seed, acc, max_bounces=depth,
                                  threads=threads, rows=rows)
            rays += int(c[0]) + int(c[1])
        rates.append(rays / (time.perf_counter() - t) / 1e6)
    return rates, rays


def _js_rates(O, view, frame, w, h, spp, depth, seed, threads, rows, trials):
    """Trials of renderImage's passes over a band by the JS restatement of the
    reference tracer on `threads` Node worker threads (after a warm-up row per
    thread).  Returns (Mrays/s per trial, rays per trial, report, sums)."""
    rep, sums = O.render_band_js(view, frame, w, h, spp, seed, rows, threads, trials,
                                 max_bounces=depth)
    rays = [t["extension"] + t["shadow"] for t in rep["trials"]]
    rates = [r / t["seconds"] / 1e6 for r, t in zip(rays, rep["trials"])]
    return rates, rays[-1], rep, sums


def _band_size(probe_s, per, budget_s, trials, h):
    """Rows (a multiple of `per`) for trials of about budget_s/trials seconds,
    from the seconds one pass set over `per` rows took."""
    return min(h, per * max(1, int(budget_s / trials / max(probe_s, 1e-3))))


def cpu_baseline(ct, ps, frame, w, h, spp, depth, seed, budget_s=12.0):
    """The reference tracer's algorithm on the host's cores, over a band of rows
    of the same frame -- same camera, seeds, depth and passes -- in the protocol
    of scripts/bench.ts:59-131: a warm-up, then three timed trials; every trial
    is reported and the best is the value.

    The value is the JS restatement of tracer.ts (oracle/js/trace.js: the
    reference's arithmetic in the JS engine, Math functions included) on Node
    worker threads, one per core the process may use, with a one-thread figure
    beside it; the C oracle's figures on the same cores follow for comparison.
    Without a node executable the C oracle is the value.  Also returns the
    band's image (renderImage's average; the JS and C restatements agree bit
    for bit) for the parity check of the timed frame."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O  # the checker, timed as the CPU baseline

    import numpy as np

    view = O.Packed(ps.prims, ps.nodes, ps.nodeCount, ps.mats, ps.lights, ps.primObject,
                    ps.order, ps.sky, getattr(ps, "leafWords", None))
    cpus = host_cpus()
    threads = cpus["allowed"]
    trials = 3
    mid = h // 2
    out = {"unit": "Mrays/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
           "host": cpus}
    acc = np.zeros(w * h * 3, np.float32)

    def c_single(rows1):
        return _band_rates(O, view, frame, w, h, spp, depth, seed, 1, (mid, mid + rows1), acc)

    def js_leg():
        # one thread: probe one row, then a band of about a second per trial
        _, _, rep, _ = _js_rates(O, view, frame, w, h, spp, depth, seed, 1, (mid, mid + 1), 1)
        rows1 = min(_band_size(rep["trials"][0]["seconds"], 1, 1.0 * trials, trials, h), h - mid)
        r1, rays1, _, _ = _js_rates(O, view, frame, w, h, spp, depth, seed, 1,
                                    (mid, mid + rows1), trials)
        # every allowed core: probe one row per thread, then the band
        y0 = max(0, mid - threads // 2)
        _, _, rep, _ = _js_rates(O, view, frame, w, h, spp, depth, seed, threads,
                                 (y0, min(h, y0 + threads)), 1)
        rows = _band_size(rep["trials"][0]["seconds"], threads, budget_s, trials, h)
        y0 = max(0, (h - rows) // 2)
        rates, rays, rep, sums = _js_rates(O, view, frame, w, h, spp, depth, seed, threads,
                                           (y0, y0 + rows), trials)
        best = max(rates)
        band = sums[y0 * w * 3:(y0 + rows) * w * 3].astype(np.float64) * (1.0 / spp)
        out.update({
            "value": best, "impl": f"JS restatement of tracer.ts on Node {rep['node']} worker "
                                   f"threads (oracle/js)",
            "trials": [round(r, 4) for r in rates],
            "single_thread": {"value": round(max(r1), 4), "unit": "Mrays/s",
                              "trials": [round(r, 4) for r in r1],
                              "sample": f"rows {mid}..{mid + rows1 - 1}, {spp} passes "
                                        f"({rays1} rays) per trial"},
            "sample": f"rows {y0}..{y0 + rows - 1} of the {w}x{h} frame, {spp} passes "
                      f"({rays} rays) per trial, {trials} trials after a warm-up row per "
                      f"thread, best {rows * w * spp * best / rays:.4f} spp*Mpix/s"})
        # the C oracle on the same cores, one trial over a band of a third of the budget
        t = time.perf_counter()
        O.accumulate_pass(view, frame, w, h, 7919 * seed, acc, max_bounces=depth,
                          threads=threads, rows=(mid, min(h, mid + threads)))
        crow = _band_size((time.perf_counter() - t) * spp, threads, budget_s / 3, 1, h)
        cy0 = max(0, (h - crow) // 2)
        cr, crays = _band_rates(O, view, frame, w, h, spp, depth, seed, threads,
                                (cy0, cy0 + crow), acc, 1)
        c1, _ = c_single(max(1, rows1 // 2))
        out["c_oracle"] = {"value": round(cr[0], 4), "cores": threads,
                           "single_thread": round(c1[0], 4),
                           "sample": f"rows {cy0}..{cy0 + crow - 1}, {spp} passes ({crays} "
                                     f"rays), one trial"}
        return out, ((y0, y0 + rows), band.astype(np.float32))

    if O.node_binary() is not None:
        try:
            return js_leg()
        except Exception as e:  # the C oracle stands in; say why
            log(f"CPU baseline: the JS leg failed ({e!r:.300}); timing the C oracle instead")
            out = {k: out[k] for k in ("unit", "cores", "kind", "cpu_model", "host")}
            out["js_error"] = repr(e)[:300]
    # no node (or the JS leg failed): the C oracle is the value
    t = time.perf_counter()
    O.accumulate_pass(view, frame, w, h, 7919 * seed, acc, max_bounces=depth, threads=1,
                      rows=(mid, mid + 1))
    rows1 = min(_band_size((time.perf_counter() - t) * spp, 1, 1.0 * trials, trials, h), h - mid)
    r1, rays1 = c_single(rows1)
    y0 = max(0, mid - threads // 2)
    t = time.perf_counter()
    O.accumulate_pass(view, frame, w, h, 7919 * seed, acc, max_bounces=depth, threads=threads,
                      rows=(y0, min(h, y0 + threads)))
    rows = _band_size((time.perf_counter() - t) * spp, threads, budget_s, trials, h)
    y0 = max(0, (h - rows) // 2)
    acc[:] = 0
    rates, rays = _band_rates(O, view, frame, w, h, spp, depth, seed, threads, (y0, y0 + rows),
                              acc, trials)
    best = max(rates)
    band = acc[y0 * w * 3:(y0 + rows) * w * 3].astype(np.float64) * (1.0 / spp)
    out.update({
        "value": best, "impl": "C oracle (" + ("the JS leg failed" if "js_error" in out
                                              else "no node executable") + ")",
        "trials": [round(r, 4) for r in rates],
        "single_thread": {"value": round(max(r1), 4), "unit": "Mrays/s",
                          "trials": [round(r, 4) for r in r1],
                          "sample": f"rows {mid}..{mid + rows1 - 1}, {spp} passes "
                                    f"({rays1} rays) per trial"},
        "sample": f"rows {y0}..{y0 + rows - 1} of the {w}x{h} frame, {spp} passes ({rays} "
                  f"rays) per trial, {trials} trials after a warm-up pass, best "
                  f"{rows * w * spp * best / rays:.4f} spp*Mpix/s"})
    return out, ((y0, y0 + rows), band.astype(np.float32))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--form", choices=["plain", "c1", "batch", "serve", "precision", "copy",
                                       "traffic"],
                    default="plain")
    ap.add_argument("--pmc-dir", default=None,
                    help="traffic: directory holding the rocprofv3 --pmc outputs in fetch/, "
                         "write/ (the plain form's runs) and copy_fetch/, copy_write/ (the "
                         "copy form's); writes the traffic profile bench.py reports")
    ap.add_argument("--copy-mib", type=int, default=1024,
                    help="copy: bytes moved by the calibration copy kernel (MiB)")
    ap.add_argument("--ref-spp", type=int, default=4096,
                    help="precision: spp of the binary64 reference image")
    ap.add_argument("--ref-cpu-pixels", type=int, default=16384,
                    help="precision: pixels of the FP64 CPU reference (the oracle on host cores, "
                         "--ref-spp samples each; 0: none)")
    ap.add_argument("--precision", choices=["f64", "f32"], default="f64",
                    help="f64: the reference's binary64 arithmetic; f32: the precision form "
                         "(quantized nodes, FP32 intersection and shading)")
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
    ap.add_argument("--persist-blocks", type=int, default=None,
                    help="workgroups of the persistent traversal launches (0: one lane per ray; "
                         "default: the tracer's, 640 binary64 / 768 FP32)")
    ap.add_argument("--rays", type=int, default=1 << 24,
                    help="plain: random closest-hit queries timed beside the render (0: none)")
    ap.add_argument("--workgroup", type=int, default=256, choices=[64, 128, 256],
                    help="lanes per workgroup of the traversal launches")
    ap.add_argument("--path-groups", type=int, default=0, choices=[0, 1, 2],
                    help="plain: path groups per round on their own streams (0: the tracer's "
                         "default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-isolated", action="store_true",
                    help="plain: skip the frame traced with every launch alone (profiling runs "
                         "that count the timed launches only)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--scenes", type=int, default=0, help="batch/serve: scene count (0: 64/128)")
    ap.add_argument("--tune-trials", type=int, default=16, help="batch: autotuner trials")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: requests per second")
    ap.add_argument("--duration", type=float, default=2.0, help="serve: seconds of arrivals")
    ap.add_argument("--rates", type=lambda v: [float(x) for x in v.split(",") if x], default=[],
                    help="serve: comma-separated rates for a latency-vs-rate curve after the "
                         "main stream (each over --curve-duration seconds, no plain comparison)")
    ap.add_argument("--curve-duration", type=float, default=10.0)
    ap.add_argument("--max-paths", type=int, default=1 << 26, help="serve: paths per round")
    ap.add_argument("--policy", choices=["fifo", "srpt", "scene"], default="scene")
    ap.add_argument("--hbm-budget-gib", type=float, default=8.0)
    ap.add_argument("--forms", type=lambda v: [x for x in v.split(",") if x],
                    default=["c1", "precision", "batch", "serve"],
                    help="plain: other configurations measured after the plain frame, summarized "
                         "in the line's 'forms' object (c1, precision, batch, serve; '' for none)")
    ap.add_argument("--forms-serve-duration", type=float, default=5.0,
                    help="plain: seconds of the request stream in the serving summary")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one process, one "
              "MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.form == "copy":
        print(json.dumps(copy_form(args)), flush=True)
        return
    if args.form == "traffic":
        print(json.dumps(traffic_form(args)), flush=True)
        return
    ct = load_package()
    ct.init(0)
    if args.form == "batch":
        res = batch_form(ct, args)
    elif args.form == "precision":
        res = precision_form(ct, args)
    elif args.form == "serve":
        res = serve_form(ct, args)
    elif args.form == "c1":
        r = c1_form(ct, args)
        res = {"metric": METRIC, "value": r.pop("value"), "unit": r.pop("unit"), "n_gpus": 1,
               "steps": r.pop("steps"), "warmup": args.warmup,
               "ms_per_step": r.pop("ms_per_step"), "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": r.pop("dtype"), "data": "synthetic",
               "config": {"workload": r.pop("workload"), "width": 256, "height": 256, "spp": 16,
                          "max_bounces": 8, "parallelism": "single GPU"},
               "roofline": None, "cpu_baseline": r.pop("cpu_baseline", None), **r}
    else:
        res = plain_form(ct, args)
        if args.forms:
            res["forms"] = other_forms(ct, args)
    print(json.dumps(res), flush=True)


def extend_roofline(cnt, prof, kernel, precision="f64", quantized=True):
    """Roofline of the closest-hit launches: algorithmic bytes from the counting
    run over the summed HIP-event time of the same launches.  Per counted node
    slot: 16 B for the 4-wide walk (a 64 B quantized node per 4 slots, both
    precisions) or 32 B for the binary walks (one float32 node each); 48 B per
    triangle and 32 B per sphere test; per ray the ray in and the hit out
    (binary64: 24 + 16 B, FP32: 24 + 8 B)."""
    launches = max(1, prof["extendLaunches"])
    node_b = 16 if quantized or precision == "f32" else 32
    ray_b = 32 if precision == "f32" else 40
    ext_bytes = (node_b * cnt["nodeVisits"] + 48 * cnt["triTests"] + 32 * cnt["sphereTests"] +
                 ray_b * cnt["extensionRays"])
    ext_s = prof["extendMs"] / 1e3
    achieved = ext_bytes / ext_s / 1e9 if ext_s > 0 else 0.0
    # VALU model of the same launches: 16 lane-ops per box test and 50 per
    # primitive test (spheres counted as triangles) over the same time
    lane_ops = 16 * cnt["nodeVisits"] + 50 * (cnt["triTests"] + cnt["sphereTests"])
    valu = lane_ops / ext_s / 1e12 if ext_s > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
            "bytes_per_launch": ext_bytes // launches,
            "avg_launch_ms": round(prof["extendMs"] / launches, 4),
            "valu_model": {"achieved": round(valu, 3), "peak": VALU_PEAK_TLOPS,
                           "unit": "T lane-ops/s", "frac": round(valu / VALU_PEAK_TLOPS, 4),
                           "lane_ops_per_launch": lane_ops // launches}}


def traversal_bytes(cnt, precision, quantized):
    """Algorithmic bytes of all traversal work of one render (extension and NEE
    rays), per the same per-node/primitive/ray figures as extend_roofline; an
    NEE ray also reads its direction, segment end and pending term (56 B)."""
    node_b = 16 if quantized or precision == "f32" else 32
    ray_b = 32 if precision == "f32" else 40
    ext = (node_b * cnt["nodeVisits"] + 48 * cnt["triTests"] + 32 * cnt["sphereTests"] +
           ray_b * cnt["extensionRays"])
    sh = (node_b * cnt["shadowNodeVisits"] + 48 * cnt["shadowTriTests"] +
          32 * cnt["shadowSphereTests"] + (ray_b + 56) * cnt["shadowRays"])
    return ext + sh


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "extend_traffic_soup1m_1080p_4spp.json")


def kernel_source_digest():
    """SHA-1 over the HIP sources and headers the library is built from (and
    the build flags), so a traffic profile can name the kernels it measured."""
    import glob
    import hashlib
    h = hashlib.sha1()
    files = sorted(glob.glob(os.path.join(ROOT, "caustic-tracer_amd", "csrc", "*")) +
                   [os.path.join(ROOT, "include", "caustic_amd.h"),
                    os.path.join(ROOT, "caustic-tracer_amd", "build.py")])
    for fn in files:
        if os.path.isfile(fn):
            h.update(os.path.basename(fn).encode())
            with open(fn, "rb") as f:
                h.update(f.read())
    return h.hexdigest()[:16]


def workload_settings(args, cfg):
    """What a traffic profile must match besides the kernel sources."""
    return {"tris": args.tris, "width": args.width, "height": args.height, "spp": args.spp,
            "depth": args.depth, "precision": args.precision, "walk": args.walk,
            "shadow_walk": args.shadow_walk, "stack": args.stack, "blocks": cfg["blocks"],
            "workgroup": cfg["workgroup"], "pathGroups": cfg["pathGroups"]}


def measured_traffic(args, cfg):
    """HBM bytes per extend launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (scaled by the copy-kernel calibration), when the profile
    was taken of these kernel sources under these settings; else None and why."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None, "no traffic profile"
    with open(TRAFFIC_PROFILE) as f:
        prof = json.load(f)
    entry = prof.get(args.precision)
    if not entry:
        return None, f"no {args.precision} entry in the traffic profile"
    if entry.get("source_digest") != kernel_source_digest():
        return None, "traffic profile taken of other kernel sources"
    if entry.get("settings") != workload_settings(args, cfg):
        return None, "traffic profile taken under other settings"
    return int(entry["traffic_bytes_per_launch"]), None


def plain_form(ct, args):
    import numpy as np
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
    # the build alone: device inputs and outputs, no allocation or copy-back timed
    import ctypes
    n = ps.primCount
    d_nodes = ct.DeviceBuffer.empty((max(1, 2 * n - 1) * 8,), np.float32)
    d_order = ct.DeviceBuffer.empty((max(1, n),), np.uint32)
    ncv = ctypes.c_uint32(0)
    dev_bt = []
    for _ in range(4):
        ct.sync()
        t = time.perf_counter()
        ct._capi.check(ct._capi.lib().ct_bvh_build_ex(
            d_b[0].ptr, d_b[1].ptr, d_b[2].ptr, n, 12, d_nodes.ptr, ctypes.byref(ncv),
            d_order.ptr, None, None), "ct_bvh_build_ex")
        ct.sync()
        dev_bt.append((time.perf_counter() - t) * 1e3)
    build_dev_ms = min(dev_bt[1:])
    del d_nodes, d_order
    log(f"GPU BVH build {build_ms:.1f} ms (best of 3, includes allocation and copy-back), "
        f"{build_dev_ms:.2f} ms on device-resident arrays")
    del d_b

    ds = ct.DeviceScene(ps)
    frame = ct.cameraFrame(scene["camera"], w / h)
    walks = {"ordered": ct.WALK_ORDERED, "stackless": ct.WALK_STACKLESS, "wide": ct.WALK_WIDE}
    walk = walks[args.walk]
    tr = ct.Tracer(walk=walk)
    tr.set_traversal(walks[args.shadow_walk], args.stack)
    if args.persist_blocks is not None:
        tr.set_schedule(args.persist_blocks > 0, max(1, args.persist_blocks))
    tr.set_precision(ct.PRECISION_F32 if args.precision == "f32" else ct.PRECISION_F64)
    tr.set_workgroup(args.workgroup)
    if args.path_groups:
        cfg = tr.get_config()
        cfg["pathGroups"] = args.path_groups
        tr.set_config(cfg)
    out = ct.DeviceBuffer.empty((w * h * 3,), np.float32)
    opts = {"width": w, "height": h, "spp": spp, "seed": 1, "maxBounces": depth,
            "russianRoulette": True, "nee": True}
    for i in range(args.warmup):
        tr.render_image(ds, frame, opts, out)
        ct.sync()
        log(f"warmup {i} done")
    ct.sync()
    # HIP events bracket every launch on its own stream during the timed steps;
    # the last step's events give the per-launch device time of each kernel
    tr.set_profiling(True)
    lat = []  # render_image returns once its round is done: per-call wall = frame latency
    t = time.perf_counter()
    for i in range(args.steps):
        t1 = time.perf_counter()
        tr.render_image(ds, frame, opts, out)
        lat.append(time.perf_counter() - t1)
    ct.sync()
    total = time.perf_counter() - t
    st = tr.stats()
    prof = st
    tr.set_profiling(False)
    rays = st["extensionRays"] + st["shadowRays"]
    ms_step = total / args.steps * 1e3
    mrays = rays * args.steps / total / 1e6
    sppmpix = w * h * spp * args.steps / total / 1e6
    log(f"timed {args.steps} steps: {ms_step:.1f} ms/step, {mrays:.0f} Mrays/s, "
        f"{sppmpix:.1f} spp*Mpix/s, rays/step ext {st['extensionRays']} shadow {st['shadowRays']}")

    # algorithmic work of the same launches (counting build of the kernels, untimed)
    tr.set_counting(True)
    tr.render_image(ds, frame, opts, out)
    cnt = tr.stats()
    tr.set_counting(False)
    kname = (f"k_extend32<{args.stack}>" if args.precision == "f32"
             else f"k_extend<{args.walk},{args.stack}>")
    cfg = tr.get_config()
    roof = extend_roofline(cnt, prof, kname, args.precision, quantized=args.walk == "wide")
    roof["traffic"], why = measured_traffic(args, cfg)
    if why:
        roof["traffic_note"] = why
    # launches of both path groups and the NEE launches run side by side, so a
    # launch's duration is a shared-chip figure; beside it, the algorithmic
    # traversal bytes of a whole step (extension + NEE rays) over the step time
    chip = traversal_bytes(cnt, args.precision, args.walk == "wide") / (ms_step / 1e3) / 1e9
    roof["concurrent_launches"] = True
    roof["step_traversal_gbs"] = round(chip, 1)
    # the same frame with every launch alone on the chip (one path group, NEE
    # launches in order on its stream): a single-kernel figure whose summed
    # launch times fit inside the frame's own time
    if not args.no_isolated:
        isolated_roofline(ct, tr, ds, frame, opts, out, cfg, roof, prof)
    log(f"extend: {prof['extendLaunches']} launches, {prof['extendMs']:.2f} ms total, "
        f"{cnt['nodeVisits'] / max(1, cnt['extensionRays']):.1f} nodes/ray, "
        f"{cnt['triTests'] / max(1, cnt['extensionRays']):.1f} tris/ray, {roof['achieved']:.0f} GB/s, "
        f"{cnt['stackFallbacks']} stack fallbacks; shade {prof['shadeMs']:.2f} ms, "
        f"shadow {prof['shadowMs']:.2f} ms "
        f"({cnt['shadowNodeVisits'] / max(1, cnt['shadowRays']):.1f} nodes/ray, "
        f"{cnt['shadowStackFallbacks']} fallbacks)")
    chq = closest_hit_queries(ct, ds, ps, args.rays) if args.rays > 0 else None
    cpu = None
    parity = None
    if not args.no_cpu:
        log("CPU baseline (the reference tracer restated in JS, Node worker threads)...")
        cpu, (rows, band) = cpu_baseline(ct, ps, frame, w, h, spp, depth, opts["seed"],
                                         args.cpu_budget)
        log(f"CPU baseline {cpu['value']:.3f} Mrays/s on {cpu['cores']} threads "
            f"({cpu['host']}), one thread {cpu['single_thread']['value']:.3f} Mrays/s "
            f"[{cpu['impl']}]; C oracle {cpu.get('c_oracle')}")
        # the timed frame against the oracle's rendering of the same rows
        g = out.numpy()[rows[0] * w * 3:rows[1] * w * 3].astype(np.float64)
        o = band.astype(np.float64)
        den = np.maximum(np.abs(g), np.abs(o))
        rel = np.where(den > 0, np.abs(g - o) / np.where(den > 0, den, 1), 0.0)
        px = rel.reshape(-1, 3).max(1)
        parity = {"reference": "oracle" if cpu["impl"].startswith("C") else
                  "oracle's JS restatement (equal to the C oracle bit for bit)",
                  "rows": [rows[0], rows[1]],
                  "pixels": int(px.size), "max_rel": float(rel.max()),
                  "pixels_over_1e-4": int((px > 1e-4).sum()),
                  "exact_fraction": round(float((g == o).mean()), 6)}
        log(f"parity of the timed frame, rows {rows[0]}..{rows[1] - 1}: max rel "
            f"{parity['max_rel']:.2e}, {parity['pixels_over_1e-4']} of {px.size} pixels over 1e-4")
    res = {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"triangle soup {args.tris} tris + 64 emissive, {w}x{h}, "
                               f"{spp} spp, maxBounces {depth}, RR+NEE, walk {args.walk}, "
                               f"precision {args.precision}, "
                               f"shadow walk {args.shadow_walk}, stack {args.stack}, "
                               f"persistent blocks {tr.get_config()['blocks']}, "
                               f"path groups {tr.get_config()['pathGroups']}",
                   "triangles": args.tris, "width": w, "height": h, "spp": spp,
                   "max_bounces": depth, "parallelism": "single GPU"},
        "spp_mpix_per_s": round(sppmpix, 3),
        "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                             "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                             "frames": len(lat)},
        "rays_per_step": rays, "bvh_build_ms": round(build_ms, 2),
        "bvh_build_device_ms": round(build_dev_ms, 2),
        "closest_hit_queries": chq,
        "roofline": roof, "cpu_baseline": cpu, "parity": parity,
    }
    return res


def c1_form(ct, args):
    """The small plain workload: the 256x256 Cornell variant, 16 spp,
    maxBounces 8, 16-bucket SAH tree built on the GPU; the CPU baseline of the
    headline line on the same frame, and the frame checked against the band of
    rows (most or all of the frame) that baseline renders with the same
    per-pixel seeds."""
    import numpy as np

    sc = ct.synth.cornell_c1()
    w = h = 256
    spp, depth = 16, 8
    ps = ct.packScene(sc, bins=16)
    ds = ct.DeviceScene(ps)
    frame = ct.cameraFrame(sc["camera"], 1.0)
    opts = {"width": w, "height": h, "spp": spp, "seed": 1, "maxBounces": depth,
            "russianRoulette": True, "nee": True}
    out = ct.DeviceBuffer.empty((w * h * 3,), np.float32)
    tr = ct.Tracer()
    steps = max(args.steps, 20)
    for _ in range(max(1, args.warmup)):
        tr.render_image(ds, frame, opts, out)
    ct.sync()
    lat = []
    t = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        tr.render_image(ds, frame, opts, out)
        lat.append(time.perf_counter() - t1)
    ct.sync()
    total = time.perf_counter() - t
    st = tr.stats()
    rays = st["extensionRays"] + st["shadowRays"]
    per_path = tr.get_path_kernel()[1] > 0
    g = out.numpy().astype(np.float64)
    res = {"workload": f"Cornell variant (scenes.ts box, light radiance 15, metal roughness 0.1, "
                       f"checker back wall; {ps.primCount} primitives, 16-bucket SAH), {w}x{h}, "
                       f"{spp} spp, maxBounces {depth}, RR+NEE",
           "value": round(rays * steps / total / 1e6, 2), "unit": "Mrays/s", "dtype": "f64",
           "steps": steps, "ms_per_step": round(total / steps * 1e3, 3),
           "spp_mpix_per_s": round(w * h * spp * steps / total / 1e6, 3),
           "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                                "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
           "rays_per_step": rays,
           "launches": "per-path (k_paths, one launch per path group)" if per_path else
                       "wavefront (extend, shade, NEE launches per bounce)"}
    if not args.no_cpu:
        # the CPU baseline of the headline line on this frame (the JS restatement on
        # worker threads, the C oracle beside it); its band, here the whole frame
        # or most of it, is the oracle's render the GPU frame is checked against
        cpu, (rows, band) = cpu_baseline(ct, ps, frame, w, h, spp, depth, opts["seed"],
                                         budget_s=6.0)
        gb = g[rows[0] * w * 3:rows[1] * w * 3]
        o = band.astype(np.float64)
        den = np.maximum(np.abs(gb), np.abs(o))
        rel = np.where(den > 0, np.abs(gb - o) / np.where(den > 0, den, 1), 0.0)
        res["parity"] = {"reference": "oracle's render of rows %d..%d" % (rows[0], rows[1] - 1),
                         "pixels": int(rel.size // 3), "max_rel": float(rel.max()),
                         "pixels_over_1e-4": int((rel.reshape(-1, 3).max(1) > 1e-4).sum()),
                         "exact_fraction": round(float((gb == o).mean()), 6)}
        res["cpu_baseline"] = cpu
    log(f"c1: {res['ms_per_step']:.2f} ms/step, {res['value']:.0f} Mrays/s, "
        f"{res['spp_mpix_per_s']:.1f} spp*Mpix/s, parity {res.get('parity')}")
    return res


# fields of each form's own line kept in the default line's "forms" object
_FORM_FIELDS = {
    "batch": ["value", "unit", "dtype", "ms_per_step", "spp_mpix_per_s", "rounds",
              "plain_sequential", "speedup_vs_plain", "identical_to_plain"],
    "precision": ["value", "unit", "dtype", "ms_per_step", "spp_mpix_per_s", "rms_vs_reference",
                  "rms_vs_fp64_cpu", "rms_same_seed_f32_vs_f64", "rms_seed_noise_f64",
                  "binary64_form", "speedup_vs_binary64_at_equal_spp"],
    "serve": ["value", "unit", "dtype", "ms_per_step", "spp_mpix_per_s", "latency_ms",
              "offered_spp_mpix_per_s", "gpu_busy_frac", "plain_one_at_a_time",
              "speedup_vs_plain", "same_as_alone"],
}


def other_forms(ct, args):
    """The other configurations, each in a shorter run than its own --form
    line (batching: 3 timed steps, no CPU leg; serving: --forms-serve-duration
    seconds of the stream), summarized for the default line."""
    out = {}
    for name in args.forms:
        sub = argparse.Namespace(**vars(args))
        t = time.perf_counter()
        log(f"form {name}...")
        try:
            if name == "c1":
                r = c1_form(ct, sub)
            elif name == "batch":
                sub.steps, sub.warmup, sub.no_cpu = 3, 1, True
                full = batch_form(ct, sub)
                r = {"workload": full["config"]["workload"],
                     **{k: full[k] for k in _FORM_FIELDS[name]},
                     "autotune": {k: full["autotune"][k] for k in ("trials", "ms", "config")},
                     "roofline": {k: full["roofline"][k] for k in ("kernel", "achieved", "frac")}}
            elif name == "precision":
                sub.steps, sub.warmup = 3, 1
                full = precision_form(ct, sub)
                r = {"workload": full["config"]["workload"],
                     **{k: full[k] for k in _FORM_FIELDS[name]},
                     "roofline": {k: full["roofline"][k] for k in ("kernel", "achieved", "frac")}}
            elif name == "serve":
                sub.duration, sub.rates = args.forms_serve_duration, []
                full = serve_form(ct, sub)
                r = {"workload": full["config"]["workload"],
                     **{k: full[k] for k in _FORM_FIELDS[name]}}
            else:
                raise ValueError(f"unknown form {name!r}")
        except Exception as e:  # one form's failure leaves the others' figures
            r = {"error": f"{type(e).__name__}: {e}"}
            log(f"form {name} failed: {r['error']}")
        r["run_s"] = round(time.perf_counter() - t, 1)
        out[name] = r
    return out


def isolated_roofline(ct, tr, ds, frame, opts, out, cfg, roof, prof):
    """The frame again with one path group and the NEE launches in order, so
    each extend launch runs alone: its average duration and the algorithmic
    bytes over it, added to roof["isolated"].  A launch alone has the chip to
    itself, so the grid is swept from the shared-chip count up to what five
    waves per SIMD fill (the occupancy of both precisions' extend kernels,
    bounded by their LDS stacks), and the fastest is kept."""
    ext_total = roof["bytes_per_launch"] * max(1, prof["extendLaunches"])
    full = 256 * 4 * 5 * 64 // cfg["workgroup"]  # CUs x SIMDs x waves x lanes / workgroup
    grids = sorted({cfg["blocks"], full * 3 // 4, full})
    sweep = []
    tr.set_overlap(False)
    for g in grids:
        tr.set_config(dict(cfg, pathGroups=1, blocks=g))
        tr.render_image(ds, frame, opts, out)  # untimed: first use of this setting
        tr.set_profiling(True)
        ct.sync()
        t = time.perf_counter()
        tr.render_image(ds, frame, opts, out)
        ct.sync()
        frame_ms = (time.perf_counter() - t) * 1e3
        iso = tr.stats()
        tr.set_profiling(False)
        sweep.append((iso["extendMs"] / max(1, iso["extendLaunches"]), g, iso, frame_ms))
    tr.set_overlap(True)
    tr.set_config(cfg)
    avg, grid, iso, iso_frame_ms = min(sweep, key=lambda x: x[0])
    iso_launches = max(1, iso["extendLaunches"])
    iso_gbs = ext_total / (iso["extendMs"] / 1e3) / 1e9 if iso["extendMs"] > 0 else 0.0
    roof["isolated"] = {
        "avg_launch_ms": round(avg, 4),
        "launches": iso["extendLaunches"], "bytes_per_launch": ext_total // iso_launches,
        "achieved": round(iso_gbs, 1), "frac": round(iso_gbs / HBM_PEAK_GBS, 4),
        "extend_ms_total": round(iso["extendMs"], 3), "frame_ms": round(iso_frame_ms, 3),
        "workgroups": grid,
        "grid_sweep": [{"workgroups": g, "avg_launch_ms": round(a, 4),
                        "frame_ms": round(f, 3)} for a, g, _, f in sweep],
        "setting": "one path group, NEE launches in order: each launch runs alone"}
    log(f"isolated extend launches: {iso['extendLaunches']} x {avg:.3f} ms = "
        f"{iso['extendMs']:.2f} ms of a {iso_frame_ms:.2f} ms frame, {iso_gbs:.0f} GB/s, "
        f"{grid} workgroups (swept {grids})")


def closest_hit_queries(ct, ds, ps, n):
    """Closest-hit-only pass on n random rays over the same scene, generated as
    bvh.test.ts:89-99 does (Rng(99): origins range(-14,14), targets
    range(-8,8), directions normalised by Math.hypot in binary64): throughput
    of the 4-wide walk, whether every ray's primitive and t equal the stackless
    walk's, and a sample checked against the oracle's ordered walk."""
    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import oracle as O  # the checker

    rays = ct.synth.bvh_test_rays(n, seed=99)
    d_rays = ct.DeviceBuffer.from_array(rays)
    prim = ct.DeviceBuffer.empty((n,), np.int32)
    tt = ct.DeviceBuffer.empty((n,), np.float64)
    uu = ct.DeviceBuffer.empty((n,), np.float64)
    vv = ct.DeviceBuffer.empty((n,), np.float64)
    lib = ct._capi.lib()

    def run():
        ct._capi.check(lib.ct_closest_hit(ds.handle, d_rays.ptr, n, ct.T_MIN, None, ct.WALK_WIDE,
                                          prim.ptr, tt.ptr, uu.ptr, vv.ptr, None),
                       "ct_closest_hit")

    run()  # warm
    ct.sync()
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        run()
    ct.sync()
    dt = (time.perf_counter() - t) / reps
    pw, tw = prim.numpy(), tt.numpy()
    s = ct.tracer.closest_hit(ds, d_rays, walk=ct.WALK_STACKLESS)
    agree = bool(np.array_equal(pw, s["prim"]) and np.array_equal(tw, s["t"]))
    k = 4096
    view = O.Packed(ps.prims, ps.nodes, ps.nodeCount, ps.mats, ps.lights, ps.primObject,
                    ps.order, ps.sky)
    o = O.closest_batch(view, rays[:k], walk=O.WALK_ORDERED)
    sample = bool(np.array_equal(pw[:k], o["prim"]) and np.array_equal(tw[:k], o["t"]))
    log(f"closest-hit queries: {n} rays, {n / dt / 1e6:.0f} Mrays/s (wide walk), all equal to "
        f"the stackless walk: {agree}, {k}-ray sample equal to the oracle: {sample}, "
        f"hits {(pw >= 0).mean():.3f}")
    return {"rays": n, "mrays_per_s": round(n / dt / 1e6, 1), "walk": "wide",
            "equal_to_stackless": agree, "oracle_sample_rays": k, "equal_to_oracle": sample,
            "hit_fraction": round(float((pw >= 0).mean()), 4)}


def _digest(a):
    import hashlib
    return hashlib.sha1(a.tobytes()).hexdigest()


def batch_form(ct, args):
    import numpy as np

    n = args.scenes or 64
    t = time.perf_counter()
    scenes, specs = ct.synth.batch_workload(n)
    tris = [sum(len(o["shape"]["indices"]) // 3 for o in sc["objects"]) for sc in scenes]
    log(f"{n} scenes generated ({sum(tris)} triangles, {min(tris)}..{max(tris)}) in "
        f"{time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    packed = [ct.packScene(sc) for sc in scenes]
    dss = [ct.DeviceScene(ps) for ps in packed]
    log(f"packed and uploaded in {time.perf_counter() - t:.1f} s")
    jobs = []
    for sp in specs:
        sc = scenes[sp["sceneIndex"]]
        j = dict(sp, scene=dss[sp["sceneIndex"]],
                 frame=ct.cameraFrame(sc["camera"], sp["width"] / sp["height"]))
        j["out"] = ct.DeviceBuffer.empty((j["width"] * j["height"] * 3,), np.float32)
        jobs.append(j)
    samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    log(f"{len(jobs)} jobs, {samples / 1e6:.1f} M paths")

    # plain: the same jobs one after another through renderImage
    plain = ct.Tracer()
    for j in jobs:
        plain.render_image(j["scene"], j["frame"], j, j["out"])
    ct.sync()
    t = time.perf_counter()
    rays_plain = 0
    for j in jobs:
        plain.render_image(j["scene"], j["frame"], j, j["out"])
        st = plain.stats()
        rays_plain += st["extensionRays"] + st["shadowRays"]
    ct.sync()
    plain_s = time.perf_counter() - t
    ref = [_digest(j["out"].numpy()) for j in jobs]
    log(f"plain sequential: {plain_s * 1e3:.1f} ms, {samples / plain_s / 1e6:.1f} spp*Mpix/s")

    tr = ct.Tracer()
    tr.set_graph(True)  # tuned as it runs: replayed graphs
    tune = tr.autotune(jobs, max_trials=args.tune_trials)
    log(f"autotune: {tune['trials']} trials, {tune['initialMs']:.1f} -> {tune['ms']:.1f} ms, "
        f"config {tune['config']}")
    for _ in range(args.warmup):
        tr.render_batch(jobs)
    ct.sync()
    t = time.perf_counter()
    for _ in range(args.steps):
        tr.render_batch(jobs)
    ct.sync()
    total = time.perf_counter() - t
    st = tr.stats()
    rays = st["extensionRays"] + st["shadowRays"]
    identical = [_digest(j["out"].numpy()) for j in jobs] == ref
    ms_step = total / args.steps * 1e3
    mrays = rays * args.steps / total / 1e6
    sppmpix = samples * args.steps / total / 1e6
    log(f"batched: {ms_step:.1f} ms/step, {mrays:.0f} Mrays/s, {sppmpix:.1f} spp*Mpix/s, "
        f"{st['rounds']} round(s), identical to plain: {identical}")
    tr.set_graph(False)
    tr.set_profiling(True)
    tr.render_batch(jobs)
    prof = tr.stats()
    tr.set_profiling(False)
    tr.set_counting(True)
    tr.render_batch(jobs)
    cnt = tr.stats()
    tr.set_counting(False)
    cfg = tune["config"]
    roof = extend_roofline(cnt, prof, f"k_extend<walk {cfg['walk']},{cfg['stackDepth']}>",
                           quantized=cfg["walk"] == ct.WALK_WIDE)
    cpu = None
    if not args.no_cpu:
        j = max(jobs, key=lambda x: x["width"])
        ps = packed[j["sceneIndex"]]
        cpu, _ = cpu_baseline(ct, ps, j["frame"], j["width"], j["height"], j["spp"],
                              j["maxBounces"], j.get("seed", 0), args.cpu_budget)
    return {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batching form: {n} soup scenes ({sum(tris)} triangles, "
                               f"{min(tris)}-{max(tris)} each), 128^2-768^2, 1-32 spp, "
                               f"maxBounces 8, one render_batch per step (autotuned, hipGraph)",
                   "scenes": n, "paths": samples, "parallelism": "single GPU"},
        "spp_mpix_per_s": round(sppmpix, 3), "rounds": st["rounds"],
        "plain_sequential": {"ms": round(plain_s * 1e3, 3),
                             "spp_mpix_per_s": round(samples / plain_s / 1e6, 3),
                             "mrays_per_s": round(rays_plain / plain_s / 1e6, 2)},
        "speedup_vs_plain": round(plain_s * 1e3 / ms_step, 3),
        "identical_to_plain": identical, "autotune": tune,
        "roofline": roof, "cpu_baseline": cpu,
    }


def serve_form(ct, args):
    import numpy as np

    n = args.scenes or 128
    t = time.perf_counter()
    counts = ct.synth.triangle_counts(n, 5, 1_000, 1_000_000, sigma=1.4)
    scenes = [ct.synth.family_scene(i, c) for i, c in enumerate(counts)]
    packed = [ct.packScene(sc) for sc in scenes]
    dss = [ct.DeviceScene(ps) for ps in packed]
    log(f"{n} resident scenes ({int(counts.sum())} triangles, {counts.min()}..{counts.max()}) "
        f"in {time.perf_counter() - t:.1f} s")
    reqs = ct.synth.request_stream(n, rate=args.rate, duration=args.duration)
    frames = [ct.cameraFrame(r["camera"], 1.0) for r in reqs]
    samples = sum(r["width"] * r["height"] * r["spp"] for r in reqs)
    log(f"{len(reqs)} requests over {args.duration} s at {args.rate}/s: {samples / 1e9:.2f} G "
        f"samples offered ({samples / args.duration / 1e6:.0f} spp*Mpix/s)")
    keep = set(range(0, len(reqs), max(1, len(reqs) // 8)))

    # plain: the same requests one at a time through renderImage, in arrival order
    plain = ct.Tracer()
    out = ct.DeviceBuffer.empty((1024 * 1024 * 3,), np.float32)
    warm = reqs[0]
    plain.render_image(dss[warm["sceneIndex"]], frames[0], dict(warm, spp=1), out)
    ct.sync()
    ref = {}
    t = time.perf_counter()
    rays_plain = 0
    per_req = np.zeros(len(reqs))  # seconds from launch to counters read (synchronous)
    req_rays = np.zeros(len(reqs))
    last = t
    for k, r in enumerate(reqs):
        t1 = time.perf_counter()
        plain.render_image(dss[r["sceneIndex"]], frames[k], r, out)
        st = plain.stats()
        per_req[k] = time.perf_counter() - t1
        req_rays[k] = st["extensionRays"] + st["shadowRays"]
        rays_plain += req_rays[k]
        if k in keep:
            ref[k] = _digest(out.numpy()[:r["width"] * r["height"] * 3])
        if time.perf_counter() - last > 10:
            last = time.perf_counter()
            log(f"plain: {k + 1}/{len(reqs)} requests in {last - t:.1f} s")
    ct.sync()
    plain_s = time.perf_counter() - t
    del plain, out
    log(f"plain one-at-a-time: {plain_s:.2f} s, {samples / plain_s / 1e6:.1f} spp*Mpix/s")
    fixed = _fixed_cost(reqs, per_req, req_rays, plain_s)
    log(f"plain fixed cost per request {fixed['fixed_ms']:.2f} ms + {fixed['ns_per_ray']:.3f} "
        f"ns/ray; requests under 2^20 samples: {fixed['small']['requests']} of {len(reqs)}, "
        f"{fixed['small']['time_frac']:.3f} of the plain time for "
        f"{fixed['small']['sample_frac']:.4f} of the samples")

    srv = ct.Server(hbm_budget=int(args.hbm_budget_gib * (1 << 30)), max_paths=args.max_paths,
                    policy=args.policy)
    wid = srv.submit(dss[0], frames[0], 64, 64, 1)  # loads the kernels
    srv.wait(wid)
    srv.release(wid)
    base = srv.stats()
    res = ct.serve.replay(srv, dss, frames, reqs, keep=keep, log=log)
    st = res["stats"]
    wall = res["wall"]
    srv.close()
    lat = res["latency"]
    same = all(_digest(res["images"][k]) == ref[k] for k in keep)
    pix = st["pixelSamples"] - base["pixelSamples"]
    rays = (st["extensionRays"] + st["shadowRays"]) - (base["extensionRays"] + base["shadowRays"])
    p50 = ct.serve.percentile(lat, 50)
    p99 = ct.serve.percentile(lat, 99)
    sppmpix = pix / wall / 1e6
    mrays = rays / wall / 1e6
    log(f"served: {wall:.2f} s wall, {sppmpix:.1f} spp*Mpix/s, {mrays:.0f} Mrays/s, p50 "
        f"{p50 * 1e3:.0f} ms, p99 {p99 * 1e3:.0f} ms, GPU busy {st['busySec'] / wall:.2f}, "
        f"{st['rounds'] - base['rounds']} rounds, same as alone: {same}")
    curve = []
    for rate in args.rates:
        curve.append(_serve_point(ct, args, dss, rate))
    return {
        "metric": METRIC, "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"serving form: {n} resident soup scenes ({int(counts.sum())} "
                               f"triangles), Poisson {args.rate}/s for {args.duration} s "
                               f"({len(reqs)} requests), 128^2-1024^2, 1-64 spp, maxBounces "
                               f"4-12, policy {args.policy}, {args.max_paths} paths per round",
                   "requests": len(reqs), "parallelism": "single GPU"},
        "spp_mpix_per_s": round(sppmpix, 3),
        "latency_ms": {"p50": round(p50 * 1e3, 2), "p99": round(p99 * 1e3, 2),
                       "max": round(max(lat) * 1e3, 2)},
        "offered_spp_mpix_per_s": round(samples / args.duration / 1e6, 3),
        "gpu_busy_frac": round(st["busySec"] / wall, 4),
        "plain_one_at_a_time": {"s": round(plain_s, 3),
                                "spp_mpix_per_s": round(samples / plain_s / 1e6, 3),
                                "mrays_per_s": round(rays_plain / plain_s / 1e6, 2)},
        "speedup_vs_plain": round(sppmpix / (samples / plain_s / 1e6), 3),
        "plain_fixed_cost": fixed, "submit_lag_ms_max": round(res["submit_lag_max"] * 1e3, 2),
        "latency_vs_rate": curve,
        "same_as_alone": same, "roofline": None, "cpu_baseline": None,
    }


def _fixed_cost(reqs, per_req, req_rays, plain_s):
    """Least-squares fit of the plain per-request time to a + b * rays: a is
    the fixed cost of a request (launches, tails, the host round trip), b the
    marginal cost of a ray; and the share of the plain time the small requests
    (under 2^20 pixel samples) take."""
    import numpy as np

    A = np.stack([np.ones_like(req_rays), req_rays], 1)
    (a, b), *_ = np.linalg.lstsq(A, per_req, rcond=None)
    samp = np.array([r["width"] * r["height"] * r["spp"] for r in reqs], np.float64)
    small = samp < (1 << 20)
    return {"fixed_ms": round(float(a) * 1e3, 3), "ns_per_ray": round(float(b) * 1e9, 4),
            "fixed_frac_of_plain": round(float(a) * len(reqs) / float(plain_s), 4),
            "median_request_ms": round(float(np.median(per_req)) * 1e3, 3),
            "small": {"requests": int(small.sum()),
                      "time_frac": round(float(per_req[small].sum() / per_req.sum()), 4),
                      "sample_frac": round(float(samp[small].sum() / samp.sum()), 5),
                      "median_ms": round(float(np.median(per_req[small])) * 1e3, 3)
                      if small.any() else None}}


def _serve_point(ct, args, dss, rate):
    """One point of the latency-vs-rate curve: the same generator at `rate`
    requests per second for `duration` seconds, served by a fresh server."""
    reqs = ct.synth.request_stream(len(dss), rate=rate, duration=args.curve_duration, seed=6)
    frames = [ct.cameraFrame(r["camera"], 1.0) for r in reqs]
    samples = sum(r["width"] * r["height"] * r["spp"] for r in reqs)
    srv = ct.Server(hbm_budget=int(args.hbm_budget_gib * (1 << 30)), max_paths=args.max_paths,
                    policy=args.policy)
    wid = srv.submit(dss[0], frames[0], 64, 64, 1)
    srv.wait(wid)
    srv.release(wid)
    base = srv.stats()
    res = ct.serve.replay(srv, dss, frames, reqs, log=log)
    st = res["stats"]
    srv.close()
    lat = res["latency"]
    pix = st["pixelSamples"] - base["pixelSamples"]
    pt = {"rate": rate, "requests": len(reqs), "wall_s": round(res["wall"], 3),
          "offered_spp_mpix_per_s": round(samples / args.curve_duration / 1e6, 2),
          "spp_mpix_per_s": round(pix / res["wall"] / 1e6, 2),
          "p50_ms": round(ct.serve.percentile(lat, 50) * 1e3, 2),
          "p99_ms": round(ct.serve.percentile(lat, 99) * 1e3, 2),
          "max_ms": round(max(lat) * 1e3, 2),
          "gpu_busy_frac": round((st["busySec"] - base["busySec"]) / res["wall"], 4),
          "rounds": st["rounds"] - base["rounds"],
          "submit_lag_ms_max": round(res["submit_lag_max"] * 1e3, 2)}
    log(f"rate {rate}/s: {pt['requests']} requests, p50 {pt['p50_ms']} ms, p99 {pt['p99_ms']} "
        f"ms, {pt['spp_mpix_per_s']} spp*Mpix/s of {pt['offered_spp_mpix_per_s']} offered, "
        f"GPU busy {pt['gpu_busy_frac']}")
    return pt


def copy_form(args):
    """A plain device-to-device copy kernel of known size (torch's elementwise
    copy), run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to calibrate how
    those counters report bytes that a kernel moves through HBM."""
    import torch

    n = args.copy_mib << 20
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    t = time.perf_counter()
    dst.copy_(src)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    return {"form": "copy", "bytes_read": n, "bytes_written": n, "launches": 4,
            "gbs": round(2 * n / dt / 1e9, 1)}


def _pmc_per_dispatch(directory, counter):
    """{kernel name: [value per dispatch]} of one counter from the
    *counter_collection.csv files rocprofv3 wrote under directory."""
    import csv
    import glob
    out = {}
    for fn in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(fn) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                out.setdefault(name, {}).setdefault(row.get("Dispatch_Id"), 0.0)
                out[name][row.get("Dispatch_Id")] += float(row["Counter_Value"])
    return {k: list(v.values()) for k, v in out.items()}


def traffic_form(args):
    """Calibrated HBM bytes per launch of the timed extend kernel, from the
    rocprofv3 FETCH_SIZE and WRITE_SIZE runs of the plain form and of the copy
    form (a copy of known size fixes how many bytes a counter unit is), stored
    with the kernel-source digest and the settings of the runs."""
    d = args.pmc_dir
    fetch = _pmc_per_dispatch(os.path.join(d, "fetch"), "FETCH_SIZE")
    write = _pmc_per_dispatch(os.path.join(d, "write"), "WRITE_SIZE")
    cf = _pmc_per_dispatch(os.path.join(d, "copy_fetch"), "FETCH_SIZE")
    cw = _pmc_per_dispatch(os.path.join(d, "copy_write"), "WRITE_SIZE")

    def pick(table, pattern):
        hits = {k: v for k, v in table.items() if pattern(k)}
        if len(hits) != 1:
            raise SystemExit(f"traffic: expected one kernel, found {sorted(hits)}")
        return next(iter(hits.items()))

    # the copy kernel moves copy_mib MiB each way per launch (its largest dispatches)
    ck, cfv = pick(cf, lambda k: "copy" in k.lower())
    _, cwv = pick(cw, lambda k: k == ck)
    nbytes = args.copy_mib << 20
    fetch_factor = nbytes / (max(cfv) * 1024)
    write_factor = nbytes / (max(cwv) * 1024)
    ext = "k_extend32<" if args.precision == "f32" else "k_extend<"
    timed = lambda k: ext in k and ", false, false," in k
    kf, fv = pick(fetch, timed)
    _, wv = pick(write, lambda k: k == kf)
    fb = sum(fv) / len(fv) * 1024 * fetch_factor
    wb = sum(wv) / len(wv) * 1024 * write_factor
    ct = load_package()
    ct.init(0)
    tr = ct.Tracer()
    tr.set_precision(ct.PRECISION_F32 if args.precision == "f32" else ct.PRECISION_F64)
    tr.set_workgroup(args.workgroup)
    if args.persist_blocks is not None:
        tr.set_schedule(args.persist_blocks > 0, max(1, args.persist_blocks))
    cfg = tr.get_config()
    if args.path_groups:
        cfg["pathGroups"] = args.path_groups
    prof = {}
    if os.path.exists(TRAFFIC_PROFILE):
        with open(TRAFFIC_PROFILE) as f:
            prof = json.load(f)
    prof["workload"] = (f"triangle soup {args.tris} tris + 64 emissive, {args.width}x"
                        f"{args.height}, {args.spp} spp, maxBounces {args.depth}")
    prof["copy_calibration"] = {"kernel": ck, "bytes_copied": nbytes,
                                "FETCH_SIZE_KiB": max(cfv), "WRITE_SIZE_KiB": max(cwv),
                                "fetch_factor": fetch_factor, "write_factor": write_factor}
    prof[args.precision] = {"kernel": kf.split("(")[0], "dispatches": len(fv),
                            "fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
                            "traffic_bytes_per_launch": fb + wb,
                            "source_digest": kernel_source_digest(),
                            "settings": workload_settings(args, cfg)}
    with open(TRAFFIC_PROFILE, "w") as f:
        json.dump(prof, f, indent=1)
    return {"form": "traffic", "precision": args.precision, **prof[args.precision]}


def precision_form(ct, args):
    import numpy as np

    t = time.perf_counter()
    sc = ct.synth.caustic_scene()
    ps = ct.packScene(sc)
    ds = ct.DeviceScene(ps)
    log(f"caustic scene: {ps.primCount} prims, {ps.nodeCount} nodes, exact leaf words "
        f"{ps.leafWords is not None}, ready in {time.perf_counter() - t:.1f} s")
    w = h = 1024
    spp, depth = 64, 16
    frame = ct.cameraFrame(sc["camera"], 1.0)
    opts = {"width": w, "height": h, "spp": spp, "seed": 1, "maxBounces": depth,
            "russianRoulette": True, "nee": True}
    out = ct.DeviceBuffer.empty((w * h * 3,), np.float32)

    ref_tr = ct.Tracer()
    t = time.perf_counter()
    ref_tr.render_image(ds, frame, dict(opts, spp=args.ref_spp, seed=1000), out)
    ct.sync()
    ref = out.numpy().astype(np.float64)
    ref_s = time.perf_counter() - t
    log(f"binary64 reference: {args.ref_spp} spp in {ref_s:.1f} s, mean {ref.mean():.5f}")
    del ref_tr

    # the FP64 CPU reference: the oracle at --ref-spp on pixels spread over the
    # image, on every core the process may use
    cpu_ref = None
    if args.ref_cpu_pixels > 0:
        sys.path.insert(0, ROOT)
        from oracle import oracle as O  # the checker, as the FP64 CPU reference

        view = O.Packed(ps.prims, ps.nodes, ps.nodeCount, ps.mats, ps.lights, ps.primObject,
                        ps.order, ps.sky, ps.leafWords)
        px = np.sort(np.random.default_rng(5).choice(w * h, args.ref_cpu_pixels, replace=False))
        threads = host_cpus()["allowed"]
        t = time.perf_counter()
        cref = O.render_pixels(view, frame, w, h, px, args.ref_spp, seed=1000,
                               max_bounces=depth, threads=threads).astype(np.float64)
        cpu_s = time.perf_counter() - t
        g = ref.reshape(-1, 3)[px]
        den = np.maximum(np.abs(g), np.abs(cref))
        rel = np.where(den > 0, np.abs(g - cref) / np.where(den > 0, den, 1), 0.0)
        cpu_ref = {"pixels": int(len(px)), "spp": args.ref_spp, "seconds": round(cpu_s, 2),
                   "threads": threads, "gpu_reference_max_rel": float(rel.max()),
                   "px": px, "values": cref}
        log(f"FP64 CPU reference: {len(px)} pixels x {args.ref_spp} spp on {threads} threads in "
            f"{cpu_s:.1f} s; the GPU binary64 reference there: max rel {rel.max():.2e}")

    res = {}
    imgs = {}
    for prec in ("f64", "f32"):
        tr = ct.Tracer()
        tr.set_precision(ct.PRECISION_F32 if prec == "f32" else ct.PRECISION_F64)
        if args.persist_blocks is not None:
            tr.set_schedule(args.persist_blocks > 0, max(1, args.persist_blocks))
        for _ in range(args.warmup):
            tr.render_image(ds, frame, opts, out)
        ct.sync()
        t = time.perf_counter()
        for _ in range(args.steps):
            tr.render_image(ds, frame, opts, out)
        ct.sync()
        total = time.perf_counter() - t
        st = tr.stats()
        img = out.numpy().astype(np.float64)
        imgs[prec] = img
        rms = float(np.sqrt(np.mean((img - ref) ** 2)))
        rays = st["extensionRays"] + st["shadowRays"]
        r = {"ms_per_step": total / args.steps * 1e3, "mrays": rays * args.steps / total / 1e6,
             "sppmpix": w * h * spp * args.steps / total / 1e6, "rms": rms,
             "rel_rms": rms / float(np.sqrt(np.mean(ref ** 2))), "rays": rays}
        tr.set_profiling(True)
        tr.render_image(ds, frame, opts, out)
        prof = tr.stats()
        tr.set_profiling(False)
        tr.set_counting(True)
        tr.render_image(ds, frame, opts, out)
        cnt = tr.stats()
        tr.set_counting(False)
        r["roofline"] = extend_roofline(cnt, prof, "k_extend32<24>" if prec == "f32" else
                                        "k_extend<wide,24>", prec)
        r["fallbacks"] = cnt["stackFallbacks"]
        if cpu_ref:
            r["rms_cpu"] = float(np.sqrt(np.mean((img.reshape(-1, 3)[cpu_ref["px"]] -
                                                  cpu_ref["values"]) ** 2)))
        log(f"{prec}: {r['ms_per_step']:.1f} ms/step, {r['mrays']:.0f} Mrays/s, "
            f"{r['sppmpix']:.1f} spp*Mpix/s, RMS vs reference {rms:.6f} "
            f"(relative {r['rel_rms']:.4f}), extend {prof['extendMs']:.1f} ms, "
            f"{cnt['nodeVisits'] / max(1, cnt['extensionRays']):.1f} nodes/ray")
        res[prec] = r
    f32, f64 = res["f32"], res["f64"]
    # arithmetic apart from sampling noise: the same seed in both precisions,
    # beside two binary64 seeds
    tr = ct.Tracer()
    tr.render_image(ds, frame, dict(opts, seed=2), out)
    ct.sync()
    seed2 = out.numpy().astype(np.float64)
    same_seed = float(np.sqrt(np.mean((imgs["f32"] - imgs["f64"]) ** 2)))
    seed_noise = float(np.sqrt(np.mean((imgs["f64"] - seed2) ** 2)))
    log(f"RMS FP32 vs binary64 at the same seed {same_seed:.3e}; binary64 seed 1 vs 2 "
        f"{seed_noise:.3e}")
    cpu_line = None
    if cpu_ref:
        cpu_line = {"pixels": cpu_ref["pixels"], "spp": cpu_ref["spp"],
                    "f32": f32["rms_cpu"], "f64": f64["rms_cpu"],
                    "gpu_binary64_reference_max_rel": cpu_ref["gpu_reference_max_rel"],
                    "cpu_seconds": cpu_ref["seconds"], "threads": cpu_ref["threads"]}
    return {
        "metric": METRIC, "value": round(f32["mrays"], 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(f32["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"precision form: {ps.primCount}-triangle caustic scene (256 "
                               f"tessellated glass spheres over a checker plane, 0.01-area "
                               f"light of radiance 500), {w}x{h}, {spp} spp, maxBounces {depth}; "
                               f"RMS against a binary64 render at {args.ref_spp} spp",
                   "triangles": ps.primCount, "width": w, "height": h, "spp": spp,
                   "max_bounces": depth, "parallelism": "single GPU"},
        "spp_mpix_per_s": round(f32["sppmpix"], 3),
        "rms_vs_reference": {"f32": f32["rms"], "f64": f64["rms"], "reference_spp": args.ref_spp,
                             "reference": "GPU binary64 render (equal per pixel to the oracle's)",
                             "f32_over_f64": round(f32["rms"] / f64["rms"], 4)},
        "rms_vs_fp64_cpu": cpu_line,
        "rms_same_seed_f32_vs_f64": same_seed, "rms_seed_noise_f64": seed_noise,
        "binary64_form": {"ms_per_step": round(f64["ms_per_step"], 3),
                          "mrays_per_s": round(f64["mrays"], 2),
                          "spp_mpix_per_s": round(f64["sppmpix"], 3),
                          "roofline": f64["roofline"]},
        "speedup_vs_binary64_at_equal_spp": round(f64["ms_per_step"] / f32["ms_per_step"], 3),
        "roofline": f32["roofline"],
        "cpu_baseline": None if not cpu_ref else {
            "value": round(cpu_ref["pixels"] * cpu_ref["spp"] / cpu_ref["seconds"] / 1e6, 4),
            "unit": "spp*Mpix/s", "cores": cpu_ref["threads"], "kind": "port",
            "sample": f"the FP64 CPU reference: {cpu_ref['pixels']} pixels x {cpu_ref['spp']} spp"},
    }


if __name__ == "__main__":
    main()
