"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``oracle/refcpu.cpp`` with g++ and OpenMP into ``oracle/librefcpu.so`` (ISA level
x86-64-v2, which any x86-64 host running this image can execute; the tests load it) and into
one library per higher ISA level, ``librefcpu_x86-64-v3.so`` (AVX2) and
``librefcpu_x86-64-v4.so`` (AVX-512).  The library is compiled in one container and run on
another host's CPU, so ``-march=native`` here would target the wrong machine; instead the
CPU baseline of bench.py loads the highest level the host it runs on supports
(refcpu.native_isa_level).  ``-ffp-contract=off`` keeps every product a separately rounded
IEEE operation, as Numba compiles the reference (no fastmath, no FMA contraction), so every
level computes the same bits.
"""

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "refcpu.cpp")
OUT = os.path.join(HERE, "librefcpu.so")

LEVELS = ("x86-64-v2", "x86-64-v3", "x86-64-v4")

FLAGS = [
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fopenmp",
    "-fPIC",
    "-shared",
]


def library_path(level: str = "x86-64-v2") -> str:
    return OUT if level == "x86-64-v2" else os.path.join(HERE, f"librefcpu_{level}.so")


def _build_one(level: str, verbose: bool) -> str:
    out = library_path(level)
    cmd = ["g++", *FLAGS, f"-march={level}", SRC, "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


def build(verbose: bool = False, levels=LEVELS) -> str:
    with cf.ThreadPoolExecutor(max_workers=len(levels)) as ex:
        list(ex.map(lambda lv: _build_one(lv, verbose), levels))
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
