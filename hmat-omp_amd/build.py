"""Build libhmat_amd.so (gfx950) in-tree with hipcc.

    python hmat-omp_amd/build.py [--jobs N]

Each HIP source is compiled to an object in hmat-omp_amd/_build/, then the
shared library is linked into hmat-omp_amd/lib/, and the C++ drop-in example
(examples/drop_in_driver.cpp) into hmat-omp_amd/bin/.  Packed FP32
instructions are disabled for every device compile.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libhmat_amd.so")
BIN_DIR = os.path.join(HERE, "bin")
DRIVER_SRC = os.path.join(HERE, "examples", "drop_in_driver.cpp")
DRIVER = os.path.join(BIN_DIR, "drop_in_driver")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["kernels.hip", "sketch.hip", "matvec.hip", "engine.hip", "hodlr.hip", "hss.hip", "capi.hip",
           "noalloc.hip", "apply.hip", "ulv.hip", "hss_mv.hip"]
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in VGPRs: no AGPR round trips around the FP32->FP64
    # flushes and fewer registers in every MFMA kernel
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-I", os.path.join(HERE, "..", "include"),
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(SRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{errs}")
    return out, errs


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs or min(len(SOURCES), 8)) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for _, w in results:
            if w.strip():
                print(w, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    # the C++ drop-in example links the library through include/hmat_amd.hpp
    os.makedirs(BIN_DIR, exist_ok=True)
    cmd = ["g++", "-O3", "-fopenmp", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"), DRIVER_SRC,
           "-L", LIB_DIR, "-lhmat_amd", "-Wl,-rpath,$ORIGIN/../lib", "-o", DRIVER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in driver build failed:\n" + r.stderr)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
