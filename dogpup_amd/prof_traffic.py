"""HBM traffic of the solver's kernel classes from two rocprofv3 PMC runs.

Usage (after ``rocprofv3 --pmc FETCH_SIZE`` and ``rocprofv3 --pmc WRITE_SIZE``
runs of the same ``bench.py`` command, each writing ``*_counter_collection.csv``
and the bench's JSON line):

    python -m dogpup_amd.prof_traffic FETCH.csv WRITE.csv --bench-json BENCH.json \
        --from-last k_fsai_ldl8 --json profiles/traffic_cocg_c64_slab269k_2048rhs.json

FETCH_SIZE and WRITE_SIZE are in KiB per dispatch; each is scaled by what it
reads for the bench's plain 2 GiB device copy (``__amd_rocclr_copyBuffer``),
so that bytes = counter x factor.  The traffic of every dispatch of a class's
kernels (0 G/G^T SpMV, 1 A SpMV, 2 Krylov vector passes, 3 FP64 refinement)
is summed over the run and divided by the algorithmic bytes the bench's
profiled step counted for that class (``extra.kernel_classes``) times the
solves the run made with that work (``--solves``).  With ``--from-last
REGEX`` only the dispatches from the last one whose kernel name matches REGEX
on are counted (``--from-last k_fsai_ldl8``: the bench's profiled step, which
is the run's last FSAI setup and solve, without the bench's own residual
checks of earlier steps), against one step's algorithmic bytes.  That
per-class ratio is what bench.py applies to its measured launches.  The
largest dispatch of every kernel is listed too.
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


def read_counter(path: str, counter: str, from_last: str | None = None):
    rows = []
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            if r["Counter_Name"] == counter:
                rows.append((int(r["Dispatch_Id"]), _short(r["Kernel_Name"]), float(r["Counter_Value"])))
    rows.sort()
    if from_last:
        marks = [d for d, n, _ in rows if re.search(from_last, n)]
        if not marks:
            raise SystemExit(f"no dispatch of a kernel matching {from_last!r} in {path}")
        rows = [r for r in rows if r[0] >= marks[-1]]
    return rows


CLASS_NAMES = ["G/G^T SpMV", "A SpMV", "Krylov vector passes", "FP64 refinement passes"]


def kernel_class(name: str):
    """Class of a complex64 solver kernel (as the bench's profiled step
    counts them), or None."""
    m = (re.match(r"dogpup::k_spmv<float, \d+, (\d), \d+(?:, \d+)?>", name)
         or re.match(r"dogpup::k_spmv_halo<\d+, (\d), \d+>", name)
         or re.match(r"dogpup::k_spmv_halo2<(\d), \d+>", name))
    if m:
        return 1 if int(m.group(1)) == 0 else 0
    if re.match(r"dogpup::k_(cg_update|update_xr|update_p|axpy_s)<float, \d+>", name):
        return 2
    if re.match(r"dogpup::k_(residual64|add_correction)<\d+>", name):
        return 3
    return None


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("fetch_csv")
    ap.add_argument("write_csv")
    ap.add_argument("--bench-json", required=True, help="the bench line of one of the two PMC runs")
    ap.add_argument("--solves", type=int, default=None,
                    help="solves of the profiled command with the same work (default 2, or 1 with --from-last)")
    ap.add_argument("--from-last", default=None,
                    help="count only the dispatches from the last one of a kernel matching this regex on")
    ap.add_argument("--copy-bytes", type=float, default=float(2 << 30))
    ap.add_argument("--json", required=True)
    a = ap.parse_args(argv)
    if a.solves is None:
        a.solves = 1 if a.from_last else 2
    # the 2 GiB copy (calibration) comes early in the run, so it is looked
    # for among all dispatches
    cf = [v for _, n, v in read_counter(a.fetch_csv, "FETCH_SIZE") if n == COPY and v > 0.4 * a.copy_bytes / 1024][:6]
    cw = [v for _, n, v in read_counter(a.write_csv, "WRITE_SIZE") if n == COPY and v > 0.4 * a.copy_bytes / 1024][:6]
    fetch = read_counter(a.fetch_csv, "FETCH_SIZE", a.from_last)
    write = read_counter(a.write_csv, "WRITE_SIZE", a.from_last)
    if not cf or not cw:
        raise SystemExit("no 2 GiB copy dispatch found for the calibration")
    f_fac = a.copy_bytes / (sum(cf) / len(cf))       # bytes per FETCH_SIZE unit
    w_fac = a.copy_bytes / (sum(cw) / len(cw))
    with open(a.bench_json) as fh:
        line = [ln for ln in fh.read().splitlines() if ln.startswith("{")][-1]
    kc = json.loads(line)["extra"]["kernel_classes"]
    kernels, classes = {}, {}
    for rows, fac, key in ((fetch, f_fac, "fetch_bytes"), (write, w_fac, "write_bytes")):
        for _, n, v in rows:
            cls = kernel_class(n)
            if cls is None:
                continue
            k = kernels.setdefault(n, {"class": cls, "dispatches": 0, "fetch_bytes": 0.0, "write_bytes": 0.0,
                                       "max_dispatch_fetch_bytes": 0.0, "max_dispatch_write_bytes": 0.0})
            if key == "fetch_bytes":
                k["dispatches"] += 1
            k[key] += v * fac
            k["max_dispatch_" + key] = max(k["max_dispatch_" + key], v * fac)
    for n, k in kernels.items():
        c = classes.setdefault(str(k["class"]), {"name": CLASS_NAMES[k["class"]], "traffic_bytes": 0.0})
        c["traffic_bytes"] += k["fetch_bytes"] + k["write_bytes"]
    for cid, c in classes.items():
        alg = kc[c["name"]]["GB"] * 1e9 * a.solves
        c["algorithmic_bytes"] = alg
        c["ratio"] = c["traffic_bytes"] / alg
    out = {"copy_calibration": {"bytes": a.copy_bytes, "fetch_factor": f_fac / 1024, "write_factor": w_fac / 1024},
           "solves": a.solves, "from_last": a.from_last, "kernels": kernels, "classes": classes}
    with open(a.json, "w") as fh:
        json.dump(out, fh, indent=1)
    print("| class | traffic GB | algorithmic GB | traffic / algorithmic |")
    print("|---|---|---|---|")
    for cid in sorted(classes):
        c = classes[cid]
        print(f"| {c['name']} | {c['traffic_bytes'] / 1e9:.1f} | {c['algorithmic_bytes'] / 1e9:.1f} | {c['ratio']:.2f} |")
    print("\n| kernel | dispatches | fetched GB | written GB | largest dispatch GB (fetch + write) |")
    print("|---|---|---|---|---|")
    for n, k in sorted(kernels.items(), key=lambda kv: -(kv[1]["fetch_bytes"] + kv[1]["write_bytes"])):
        print(f"| `{n}` | {k['dispatches']} | {k['fetch_bytes'] / 1e9:.1f} | {k['write_bytes'] / 1e9:.1f} | "
              f"{(k['max_dispatch_fetch_bytes'] + k['max_dispatch_write_bytes']) / 1e9:.2f} |")
    print(f"\ncalibration: bytes = FETCH_SIZE x {f_fac / 1024:.3f} KiB, WRITE_SIZE x {w_fac / 1024:.3f} KiB")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
