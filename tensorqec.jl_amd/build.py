"""Build the HIP library ``tensorqec/libtqec_hip.so`` for gfx950, in tree.

Usage: ``python tensorqec.jl_amd/build.py`` (or ``__graft_entry__.build()``).
hipcc cross-compiles without a GPU.  ``python tensorqec.jl_amd/build.py
--sanitize`` builds ``tensorqec/libtqec_hip_san.so`` instead: the same
library with its host code (descriptor parser, plan compiler, kernel source
generator, C ABI) under AddressSanitizer and UndefinedBehaviorSanitizer, for
the CPU tests; it needs the ASan runtime preloaded (``sanitizer_runtime()``).
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
OUT_SAN = os.path.join(HERE, "tensorqec", "libtqec_hip_san.so")

SOURCES = ["tqec_hip.hip", "plan_compile.cpp", "spec_codegen.cpp", "mlp_hip.hip", "sampler_hip.hip"]

# Packed FP32 instructions are disabled for all device code: the compiler is
# free to form them from float2 arithmetic otherwise.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# Sanitizers for the host side only: every -fsanitize= behind -Xarch_host.
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
             "-Xarch_host", "-g"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sanitizer_runtime() -> str:
    """The shared ASan runtime of hipcc's clang (to preload into Python)."""
    import glob
    for c in sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")):
        return c
    raise RuntimeError("libclang_rt.asan-x86_64.so not found")


def build(verbose: bool = False, extra=None, sanitize: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    objs = []
    bdir = os.path.join(HERE, "_build", "san") if sanitize else os.path.join(HERE, "_build")
    extra = list(extra or []) + (SAN_FLAGS if sanitize else [])
    os.makedirs(bdir, exist_ok=True)
    for s in srcs:
        o = os.path.join(bdir, os.path.basename(s) + ".o")
        cmd = [hipcc(), "-x", "hip", "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-c", s, "-o", o,
               "-I", INCLUDE, "-I", CSRC, "-Wall", "-Wno-unused-function"] + DEVICE_FLAGS + list(extra or [])
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(o)
    out = OUT_SAN if sanitize else OUT
    tmp = out + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lhiprtc"]
    if sanitize:
        cmd += ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libsan"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, sanitize="--sanitize" in sys.argv))
