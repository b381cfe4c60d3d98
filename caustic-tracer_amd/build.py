"""Builds libcaustic_amd.so (gfx950) in-tree with hipcc.

Flags that matter for results:
  -ffp-contract=off           no a*b+c fusion: each binary64 op rounds once, as in V8
  -fno-gpu-flush-denormals-to-zero
  -target-feature -packed-fp32-ops   no v_pk_*_f32 generation
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcaustic_amd.so")
SOURCES = ["ct_runtime.hip", "ct_trace.hip", "ct_render.hip", "ct_render32.hip", "ct_bvh.hip",
           "ct_tune.hip", "ct_serve.hip", "ct_display.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
OBJ_DIR = os.path.join(HERE, "_build")


def _compile(src, obj, verbose):
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose=False):
    """Compiles every translation unit (in parallel) and links the library."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = [os.path.join(OBJ_DIR, os.path.splitext(os.path.basename(s))[0] + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=max(1, min(len(srcs), os.cpu_count() or 1))) as ex:
        results = list(ex.map(lambda so: _compile(so[0], so[1], verbose), zip(srcs, objs)))
    failed = [(s, r) for s, r in zip(srcs, results) if r.returncode != 0]
    for s, r in failed:
        sys.stderr.write(f"--- {os.path.basename(s)}\n" + r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed building libcaustic_amd.so")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libcaustic_amd.so")
    os.replace(tmp, LIB)
    return LIB


NODE_DIR = os.path.join(HERE, "node")
ADDON = os.path.join(NODE_DIR, "caustic_amd.node")
NODE_INCLUDE = "/usr/include/node"


def build_addon(verbose=False):
    """N-API addon (plain shared object against node_api.h, linked to the library)."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-I", NODE_INCLUDE,
           os.path.join(NODE_DIR, "caustic_napi.c"), "-o", ADDON + ".tmp", "-L", HERE,
           "-lcaustic_amd", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("gcc failed building the N-API addon")
    os.replace(ADDON + ".tmp", ADDON)
    return ADDON


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_addon(verbose=True))
