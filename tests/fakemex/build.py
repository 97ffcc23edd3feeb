"""Build the fake-mex harness for the gateway tests (test infrastructure).

    python tests/fakemex/build.py

Outputs under tests/fakemex/lib/:
  libfakemex.so          the mxArray store and error path (fakemex.cpp)
  librecord_dogpup.so    a recording stand-in for the mex-level entry points
  <gateway>_rec.so       each gateway of dogpup_amd/mex/ linked against the stand-in
  <gateway>.so           each gateway linked against dogpup_amd/lib/libdogpup_hip.so
                         (only when that library is built)
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
MEX = ROOT / "dogpup_amd" / "mex"
OUT = HERE / "lib"
GATEWAYS = ("smv_mex", "FSAImv_mex", "FSAIP_gen_mex")
FLAGS = ["-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Werror"]


def _cc(args):
    cmd = [os.environ.get("CXX", "g++"), *FLAGS, *args]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"fake-mex build failed: {' '.join(cmd)}\n{r.stderr}")


def _sources():
    return [HERE / "fakemex.cpp", HERE / "mex.h", HERE / "record_dogpup.cpp", ROOT / "include" / "dogpup.h",
            *sorted(MEX.glob("*.cpp")), *sorted(MEX.glob("*.hpp"))]


def up_to_date() -> bool:
    """every output exists and is newer than every source (the gateways
    linked against libdogpup_hip.so also newer than that library)"""
    outs = [OUT / "libfakemex.so", OUT / "librecord_dogpup.so", *(OUT / f"{g}_rec.so" for g in GATEWAYS),
            *(OUT / f"{g}.so" for g in GATEWAYS)]
    if not all(o.exists() for o in outs):
        return False
    newest = max(s.stat().st_mtime for s in _sources())
    real = ROOT / "dogpup_amd" / "lib" / "libdogpup_hip.so"
    if real.exists():
        newest = max(newest, real.stat().st_mtime)
    return min(o.stat().st_mtime for o in outs) >= newest


def build(verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    inc = ["-I", str(HERE), "-I", str(ROOT / "include")]
    _cc([*inc, str(HERE / "fakemex.cpp"), "-o", str(OUT / "libfakemex.so")])
    _cc([*inc, str(HERE / "record_dogpup.cpp"), "-o", str(OUT / "librecord_dogpup.so")])
    real = ROOT / "dogpup_amd" / "lib" / "libdogpup_hip.so"
    for g in GATEWAYS:
        src = str(MEX / f"{g}.cpp")
        _cc([*inc, src, "-o", str(OUT / f"{g}_rec.so"), "-L", str(OUT), "-lfakemex", "-lrecord_dogpup",
             "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined"])
        if real.exists():
            _cc([*inc, src, "-o", str(OUT / f"{g}.so"), "-L", str(OUT), "-lfakemex", "-L", str(real.parent),
                 "-ldogpup_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/../../../dogpup_amd/lib",
                 "-Wl,--no-undefined"])
    if verbose:
        print(f"built the fake-mex gateways in {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
