"""Build libtsphere.so (the gfx950 HIP library behind include/tsphere.h).

Run as ``python turing-sphere_amd/build.py``; ``__graft_entry__.build()``
calls :func:`build`.  The library is built in-tree so it travels with the
source tree; every HIP translation unit is compiled for gfx950 only, with
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtsphere.so")
SOURCES = ["tsphere_runtime.hip", "sht_legendre.hip", "sht_legendre_mfma.hip",
           "sht_fourier.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


# Per-source additions.  The Legendre kernels keep their matrix-core
# accumulators in VGPRs: with the default AGPR form the compiler rotated the
# 4x4-block accumulators through accvgpr moves inside the l-loop (24 moves per
# 32 VALU ops of the recurrence).
EXTRA = {"sht_legendre.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
OBJDIR = os.path.join(HERE, "_build")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature: harmless.
    noise = "'-packed-fp32-ops' is not a recognized feature"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)


def build(verbose: bool = False) -> str:
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *FLAGS, "-c", *EXTRA.get(src, []), "-I", os.path.join(ROOT, "include"),
                     "-I", CSRC, os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = OUT + ".tmp"
    _run([HIPCC, *FLAGS, "-shared", *objs, "-o", tmp, "-lhiprtc"], verbose)
    os.replace(tmp, OUT)
    build_napi(verbose)
    return OUT


NODE_INCLUDE = os.environ.get("NODE_INCLUDE", "/usr/include/node")
NAPI_OUT = os.path.join(HERE, "napi", "tsphere_napi.node")


def build_napi(verbose: bool = False):
    """The N-API addon (napi/tsphere_napi.cc) for a node host, when node's
    headers are installed; returns its path, or None when they are not."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print(f"node headers not found under {NODE_INCLUDE}: N-API addon not built",
                  file=sys.stderr)
        return None
    tmp = NAPI_OUT + ".tmp"
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-Wall", "-I", NODE_INCLUDE,
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "napi", "tsphere_napi.cc"),
           "-L", HERE, "-ltsphere", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"N-API addon build failed ({res.returncode}):\n{res.stderr}")
    os.replace(tmp, NAPI_OUT)
    return NAPI_OUT


if __name__ == "__main__":
    print(build(verbose=True))
