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
SOURCES = ["ct_runtime.hip", "ct_trace.hip", "ct_render.hip", "ct_bvh.hip", "ct_tune.hip",
           "ct_serve.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libcaustic_amd.so")
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
