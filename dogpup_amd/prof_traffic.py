"""HBM traffic per kernel launch from two rocprofv3 PMC runs.

Usage (after ``rocprofv3 --pmc FETCH_SIZE`` and ``rocprofv3 --pmc WRITE_SIZE``
runs of the same ``bench.py`` command, each writing ``*_counter_collection.csv``):

    python -m dogpup_amd.prof_traffic FETCH.csv WRITE.csv --method cocg \
        --json profiles/traffic_cocg_c64_slab269k_2048rhs.json

FETCH_SIZE and WRITE_SIZE are in KiB per dispatch; each is scaled by what it
reads for the bench's plain 2 GiB device copy (``__amd_rocclr_copyBuffer``),
so that bytes = counter x factor.  For every solver kernel the dispatch with
the largest traffic is taken (all frequency groups active) and compared with
the algorithmic bytes of such a launch at the workload's sizes.  The
per-class ratios (traffic / algorithmic summed over the kernels of a class:
0 G/G^T SpMV, 1 A SpMV, 2 Krylov vector passes, 3 FP64 refinement) are what
bench.py applies to its measured launches.
"""
from __future__ import annotations

import argparse
import csv
import json
import re

COPY = "__amd_rocclr_copyBuffer"


def _short(name: str) -> str:
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    depth = 0
    for i, ch in enumerate(name):      # cut the parameter list after the template
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def read_counter(path: str, counter: str):
    rows = []
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            if r["Counter_Name"] == counter:
                rows.append((int(r["Dispatch_Id"]), _short(r["Kernel_Name"]), float(r["Counter_Value"])))
    return rows


def algorithmic(name: str, method: str, nodes: int, groups: int, nsp: int, nnz_a: int, nnz_g: int):
    """(class, algorithmic bytes) of a full-activity launch of a complex64
    solver kernel, or None for kernels outside the table."""
    V = float(nodes) * nsp * 8 * groups              # one complex64 vector stream, all groups
    MA = nnz_a * 12.0 + 4.0 * (nodes + 1)            # A: re, im (float) and column, once per launch
    MG = groups * nnz_g * 12.0 + 4.0 * (nodes + 1)   # G: complex64 per frequency and column
    m = (re.match(r"dogpup::k_spmv<float, (\d+), (\d), (\d)(?:, \d+)?>", name)
         or re.match(r"dogpup::k_spmv_halo<(\d+), (\d), (\d)>", name)
         or re.match(r"dogpup::k_spmv_halo2<()(\d), (\d)>", name))
    if m:
        mat, epi = int(m.group(2)), int(m.group(3))
        # vector streams per epilogue (STORE_DOT in COCG reads w = x, the
        # gathered vector); AXPY_SELF: x in, y in and out; STORE_SELF2: x, y
        streams = {0: 2, 1: 2 if method == "cocg" else 3, 2: 2, 3: 2, 4: 3, 5: 5, 6: 3, 7: 2}[epi]
        return (1, streams * V + MA) if mat == 0 else (0, streams * V + MG)
    table = {f"dogpup::k_cg_update<float, {nsp}>": (2, 3 * V), f"dogpup::k_update_xr<float, {nsp}>": (2, 8 * V),
             f"dogpup::k_update_p<float, {nsp}>": (2, 4 * V), f"dogpup::k_axpy_s<float, {nsp}>": (2, 3 * V),
             # x64 and b in complex128 (2 V each), r and x out in complex64 (+ p and rhat for
             # BiCGSTAB), A in FP64
             f"dogpup::k_residual64<{nsp}>": (3, (6 if method == "cocg" else 8) * V + nnz_a * 20.0
                                              + 4.0 * (nodes + 1))}
    return table.get(name)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch_csv")
    ap.add_argument("write_csv")
    ap.add_argument("--method", default="cocg", choices=["cocg", "bicgstab"])
    ap.add_argument("--nodes", type=int, default=269001)
    ap.add_argument("--groups", type=int, default=64, help="frequency x column-block groups")
    ap.add_argument("--nsp", type=int, default=32)
    ap.add_argument("--nnz-a", type=int, default=3930201)
    ap.add_argument("--nnz-g", type=int, default=2099601)
    ap.add_argument("--copy-bytes", type=float, default=float(2 << 30))
    ap.add_argument("--json", required=True)
    a = ap.parse_args(argv)
    fetch = read_counter(a.fetch_csv, "FETCH_SIZE")
    write = read_counter(a.write_csv, "WRITE_SIZE")
    cf = [v for _, n, v in fetch if n == COPY and v > 0.4 * a.copy_bytes / 1024][:6]
    cw = [v for _, n, v in write if n == COPY and v > 0.4 * a.copy_bytes / 1024][:6]
    if not cf or not cw:
        raise SystemExit("no 2 GiB copy dispatch found for the calibration")
    f_fac = a.copy_bytes / (sum(cf) / len(cf))       # bytes per FETCH_SIZE unit
    w_fac = a.copy_bytes / (sum(cw) / len(cw))
    wmax = {}
    for _, n, v in write:
        wmax.setdefault(n, []).append(v)
    kernels, classes = {}, {}
    for n in sorted({n for _, n, _ in fetch}):
        alg = algorithmic(n, a.method, a.nodes, a.groups, a.nsp, a.nnz_a, a.nnz_g)
        if alg is None:
            continue
        cls, alg_b = alg
        fb = max(v for _, m, v in fetch if m == n) * f_fac
        wb = max(wmax.get(n, [0.0])) * w_fac
        kernels[n] = {"class": cls, "fetch_bytes": fb, "write_bytes": wb, "traffic_bytes": fb + wb,
                      "algorithmic_bytes": alg_b, "ratio": (fb + wb) / alg_b}
        c = classes.setdefault(str(cls), {"traffic_bytes": 0.0, "algorithmic_bytes": 0.0})
        c["traffic_bytes"] += fb + wb
        c["algorithmic_bytes"] += alg_b
    for c in classes.values():
        c["ratio"] = c["traffic_bytes"] / c["algorithmic_bytes"]
    out = {"method": a.method, "copy_calibration": {"bytes": a.copy_bytes, "fetch_factor": f_fac / 1024,
                                                    "write_factor": w_fac / 1024},
           "kernels": kernels, "classes": classes}
    with open(a.json, "w") as fh:
        json.dump(out, fh, indent=1)
    print("| kernel | fetched GB | written GB | traffic GB | algorithmic GB | traffic / algorithmic |")
    print("|---|---|---|---|---|---|")
    for n, k in kernels.items():
        print(f"| `{n}` | {k['fetch_bytes'] / 1e9:.2f} | {k['write_bytes'] / 1e9:.2f} | "
              f"{k['traffic_bytes'] / 1e9:.2f} | {k['algorithmic_bytes'] / 1e9:.2f} | {k['ratio']:.2f} |")
    print(f"\ncalibration: bytes = FETCH_SIZE x {f_fac / 1024:.3f} KiB, WRITE_SIZE x {w_fac / 1024:.3f} KiB")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
