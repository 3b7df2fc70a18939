"""Build recipe for the gfx950 library, its CLI and the tests' oracle.

Outputs (in-tree and git-ignored; they travel with the source tree):
  rtnn_amd/librtnn_hip.so   C ABI of include/rtnn_hip.h and include/rtnn_data.h
  rtnn_amd/rtnn             optixNSearch-compatible command-line driver

Every HIP translation unit is compiled for gfx950 only, with
  -ffp-contract=off   d2 = ((dx*dx + dy*dy) + dz*dz) stays five rounded FP32
                      operations on device and host (no FMA), as in the oracle;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librtnn_hip.so")
CLI = os.path.join(HERE, "rtnn")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-I", os.path.join(ROOT, "include")]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-fno-gpu-rdc"]

SOURCES = ["sort.hip", "grid.hip", "search.hip", "partition.hip", "mlp.hip", "api.hip", "data.cpp"]
# Per-file flags.  mlp.hip: MFMA results straight into VGPRs (the GELU reads
# every accumulator; in AGPRs each would first need a v_accvgpr_read).
EXTRA = {"mlp.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def _compile(src, verbose):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, *COMMON, *EXTRA.get(src, []), "-x", "hip", "-c", path, "-o", obj]
    else:
        cmd = ["g++", *COMMON, "-c", path, "-o", obj]
    out = _run(cmd, verbose)
    return obj, out


def build(verbose: bool = False, jobs: int = 6) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    objs = [o for o, _ in results]
    for _, out in results:
        if out.strip() and verbose:
            print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    cli_src = os.path.join(CSRC, "cli.cpp")
    if os.path.exists(cli_src):
        tmp = CLI + ".tmp"
        knn = os.environ.get("RTNN_KNN", "50")  # compile-time K of KNN mode
        _run(["g++", *COMMON, f"-DRTNN_KNN={int(knn)}", cli_src, "-o", tmp, "-L", HERE,
              "-lrtnn_hip",
              f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"], verbose)
        os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
