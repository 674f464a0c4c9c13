"""Builds the MI355X tick engine, libd3fgpu.so, from csrc/*.hip with hipcc
(gfx950), and the N-API addon when Node's headers are present.

Floating point flags: -ffp-contract=off (no FMA contraction: every FP64
operation rounds once, as the reference's wasm core does) and no fast-math.
Packed FP32 instructions are disabled (-packed-fp32-ops).
"""

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libd3fgpu.so")
SOURCES = ["runtime.hip", "tree.hip", "many_body.hip", "link.hip", "collide.hip", "simple.hip", "fast.hip", "fmm.hip", "batch.hip", "batch_exact.hip", "gnn.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIPFLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stdout + r.stderr))
    warnings = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                         if "packed-fp32-ops" not in l and l.strip())
    return obj, warnings


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        results = list(pool.map(_compile, SOURCES))
    if verbose:
        for obj, w in results:
            if w:
                print(w, file=sys.stderr)
    tmp = LIBRARY + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in results],
           "-Wl,-Bsymbolic"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    os.replace(tmp, LIBRARY)
    return LIBRARY


def build_addon(verbose=False):
    """The N-API addon (napi/addon.cc) over libd3fgpu.so, when node headers exist."""
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")) or shutil.which("g++") is None:
        return None
    out = os.path.join(HERE, "js", "d3fgpu.node")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", inc, "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "napi", "addon.cc"), "-o", out, "-L", HERE, "-ld3fgpu",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("addon build failed:\n%s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    return out


def build(verbose=False):
    lib = build_library(verbose)
    addon = None
    if os.path.exists(os.path.join(HERE, "napi", "addon.cc")):
        addon = build_addon(verbose)
    return lib, addon


if __name__ == "__main__":
    print(build(verbose=True))
