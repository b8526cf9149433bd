"""Build the HIP library ``tensorqec/libtqec_hip.so`` for gfx950, in tree.

Usage: ``python tensorqec.jl_amd/build.py`` (or ``__graft_entry__.build()``).
hipcc cross-compiles without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(REPO, "include")
OUT = os.path.join(HERE, "tensorqec", "libtqec_hip.so")

SOURCES = ["tqec_hip.hip", "plan_compile.cpp", "spec_codegen.cpp", "mlp_hip.hip", "sampler_hip.hip"]

# Packed FP32 instructions are disabled for all device code: the compiler is
# free to form them from float2 arithmetic otherwise.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, extra=None) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for s in srcs:
        o = os.path.join(bdir, os.path.basename(s) + ".o")
        cmd = [hipcc(), "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-c", s, "-o", o,
               "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function"] + DEVICE_FLAGS + list(extra or [])
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(o)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lhiprtc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
