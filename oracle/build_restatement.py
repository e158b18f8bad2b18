"""Compile the C++/OpenMP restatement oracle/hodlr_omp.cpp (test
infrastructure: the CPU checker and the CPU baseline, never the product).

    python oracle/build_restatement.py      ->  oracle/bin/hodlr_omp

g++ only; BLAS/LAPACK is resolved by the program at run time (dlopen), so
the link needs nothing beyond libdl.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hodlr_omp.cpp")
EXE = os.path.join(HERE, "bin", "hodlr_omp")


def build():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    tmp = EXE + ".tmp"
    cmd = ["g++", "-O3", "-fopenmp", "-std=c++17", "-Wall", SRC, "-o", tmp, "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hodlr_omp build failed:\n" + r.stderr)
    os.replace(tmp, EXE)
    return EXE


def ensure():
    """The built program; compiled first when it is missing or older than its source."""
    if not os.path.exists(EXE) or os.path.getmtime(EXE) < os.path.getmtime(SRC):
        build()
    return EXE


if __name__ == "__main__":
    print(build())
