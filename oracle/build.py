"""Builds the CPU oracle (oracle/liboracle.so) with g++ -fopenmp.

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's CPU
baseline load it as the checker.  The reference sources cannot be compiled here
(they need nvcc, CUDA SDK helper headers and CUDPP), so there is no oracle/_ref.
"""
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build(verbose: bool = False) -> pathlib.Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
           str(HERE / "oracle.cpp"), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
