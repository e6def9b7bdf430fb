"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``oracle/refcpu.cpp`` into ``oracle/librefcpu.so`` with g++ and OpenMP.
``-ffp-contract=off`` keeps every product a separately rounded IEEE operation, as
Numba compiles the reference (no fastmath, no FMA contraction).  The target ISA is
x86-64-v2 so the library also runs on a host CPU other than the build machine's.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "refcpu.cpp")
OUT = os.path.join(HERE, "librefcpu.so")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-march=x86-64-v2",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fopenmp",
    "-fPIC",
    "-shared",
]


def build(verbose: bool = False) -> str:
    cmd = ["g++", *FLAGS, SRC, "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
