"""Build oracle/lib/liboracle_cpu.so, the OpenMP FP64 restatement of the
FSAI setup and the reference-structure BiCGSTAB (oracle/cpu_bicgstab.cpp).

    python -m oracle.build_cpu

Test infrastructure only (the timed CPU baseline and a checker)."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "cpu_bicgstab.cpp"
LIB = HERE / "lib" / "liboracle_cpu.so"
# x86-64-v3 (AVX2, FMA): every current server x86 host
FLAGS = ["-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-march=x86-64-v3", "-mtune=generic"]


def build(verbose: bool = False) -> Path:
    LIB.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [os.environ.get("CXX", "g++"), *FLAGS, str(SRC), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle CPU build failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
