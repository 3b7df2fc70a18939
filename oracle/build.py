"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/oracle.c into oracle/liboracle.so with gcc.  The flags keep
FP32 arithmetic IEEE round-to-nearest with no contraction, so the oracle's d2 is
the reference's FP32 formula evaluated operation by operation.

The reference's own search path cannot be compiled here (it is CUDA + OptiX 7.1
device programs compiled at run time by NVRTC, and its headers need CUDA's
vector_functions.h), so oracle/_ref/ stays empty: there is no reference build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "oracle.c")
OUT = os.path.join(HERE, "liboracle.so")

CFLAGS = ["-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
          "-shared", "-pthread", "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False) -> str:
    cmd = ["gcc", *CFLAGS, SRC, "-o", OUT + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
