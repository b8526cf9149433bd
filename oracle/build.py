"""Build the oracle's C executor (oracle/_build/libtn_oracle.so) with gcc.

TEST INFRASTRUCTURE ONLY.  The reference (TensorQEC.jl) is pure Julia, so
there is no reference C/C++ source to compile into oracle/_ref.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "libtn_oracle.so")


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall",
           os.path.join(HERE, "cexec.c"), "-o", OUT + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(True))
