"""Autotuner for the batched solver's kernel shapes.

For a given set of meshes (one mesh or a MeshBatch), it times a few capped
Krylov iterations of the real problem for each candidate:
- group width NSP, the RHS block of the SpMV: 8, 16 or 32 columns per group.
  Narrow groups waste no padding columns on meshes with few sources; wide
  groups read each matrix entry once for more RHS.
- row tile of the SpMV and vector passes: 256, 512 or 1024 rows.
- x gathers in flight per row in the A and G SpMVs: 4 or 8.

It keeps the fastest per RHS-iteration.  With ``strides`` (mixed mode) it then
times whole solves of the chosen shape at ``tol`` with Galerkin frequency
continuation for each level-0 spacing of the schedule (every stride-th
frequency solved from zero): a wider spacing solves fewer frequencies from
zero but runs more levels, each a synchronised phase with fewer groups, and
the best spacing depends on the meshes' frequency counts. With ``splits`` it
also times whole solves of SplitBatch partitions (the meshes with more than a
threshold of columns in wider groups, each part its own handle, solved one
after the other) against the single handle, and keeps the fastest. Results
are cached in a JSON file
(``$DOGPUP_TUNE_CACHE``, default ``~/.cache/dogpup_amd/tune.json``) under a
signature of the problem: node, frequency and column counts per mesh, mode and
method.
"""
from __future__ import annotations

import gc
import json
import os
import time

import torch

from . import _lib
from .batch import MeshBatch, SplitBatch

TILES = (256, 512, 1024)
UNROLLS = (4, 8)
WIDTHS = (8, 16, 32)
STRIDES = (8, 16, 32)


def _cache_path():
    return os.environ.get("DOGPUP_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "dogpup_amd", "tune.json"))


def signature(meshes, which, mode, method):
    cols = [sum(len((m.optode.s_bary if w == "s" else m.optode.d_bary)[0]) for w in which) for m in meshes]
    return json.dumps({"nn": [m.nn for m in meshes], "nf": [m.optode.Nf for m in meshes], "ncols": cols,
                       "mode": mode, "method": method})


def _time_iterations(B, iters):
    """Seconds per Krylov iteration of the whole batch (a capped solve with an
    unreachable tolerance, after one warm-up that captures the graph)."""
    b = B.rhs()
    B.solve(b, tol=1e-30, maxit=iters)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.solve(b, tol=1e-30, maxit=iters)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del b
    return dt / iters


def _time_solve(B, tol):
    """Seconds of a whole solve (after one that builds the schedule and
    captures the graphs)."""
    b = B.rhs()
    B.solve(b, tol=tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    B.solve(b, tol=tol)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del b
    return dt


def _ncols(meshes, which):
    return [sum(len((m.optode.s_bary if w == "s" else m.optode.d_bary)[0]) for w in which) for m in meshes]


def tune(meshes, which="s", mode="mixed", method="cocg", iters=6, widths=WIDTHS, tiles=TILES, unrolls=UNROLLS,
         strides=None, splits=None, tol=5e-8, use_cache=True, verbose=False):
    """Returns ({"nsp", "tile_rows", "unroll", "s_per_iter"[, "cont_stride",
    "s_per_solve"][, "split"]}, table of every candidate: seconds per
    iteration, and seconds per solve for the continuation spacings and the
    partitions).  ``splits``: (wide, threshold) pairs, SplitBatch.width_parts
    with the tuned width as the narrow one; "split" is kept when one of them
    solves faster than the single handle."""
    meshes = list(meshes)
    key = signature(meshes, which, mode, method)
    if strides and mode == "mixed":
        key += json.dumps({"strides": list(strides), "tol": tol})
    if splits:
        key += json.dumps({"splits": [list(x) for x in splits], "tol": tol})
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]["best"], cache[key]["table"]
    maxcols = max(sum(len((m.optode.s_bary if w == "s" else m.optode.d_bary)[0]) for w in which)
                  for m in meshes)
    table = []
    best = None
    for nsp in widths:
        if nsp > 8 and nsp // 2 >= maxcols:   # wider than twice the widest mesh: pure padding
            continue
        B = None
        try:
            # every RHS from 0 while timing: the per-iteration cost, not the
            # continuation's iteration savings, is what the shapes change
            B = MeshBatch(meshes, which, mode=mode, nsp=nsp, method=method, cont=False)
            B.ensure_matrix()
            for tile in tiles:
                for u in unrolls:
                    B.set_tuning(tile, u, u, 0)
                    s = _time_iterations(B, iters)
                    row = {"nsp": nsp, "tile_rows": tile, "unroll": u, "s_per_iter": s}
                    table.append(row)
                    if verbose:
                        print(f"[tune] nsp {nsp:2d} tile {tile:4d} unroll {u:2d}: {s * 1e3:.2f} ms/iteration",
                              flush=True)
                    if best is None or s < best["s_per_iter"]:
                        best = dict(row)
        except (torch.OutOfMemoryError, _lib.DogpupError) as e:
            # a width whose padded vectors do not fit is not a candidate
            if isinstance(e, _lib.DogpupError) and e.code != _lib.ERR_NOMEM:
                raise
            if verbose:
                print(f"[tune] nsp {nsp:2d}: does not fit in device memory, skipped", flush=True)
        del B
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    if best is None:
        raise RuntimeError("no candidate shape fits in device memory")
    if strides and mode == "mixed":
        B = MeshBatch(meshes, which, mode=mode, nsp=best["nsp"], method=method,
                      tuning=(best["tile_rows"], best["unroll"], best["unroll"], 0))
        for st in strides:
            B.set_cont_stride(st)
            s = _time_solve(B, tol)
            table.append({"nsp": best["nsp"], "tile_rows": best["tile_rows"], "unroll": best["unroll"],
                          "cont_stride": st, "s_per_solve": s})
            if verbose:
                print(f"[tune] continuation stride {st:2d}: {s * 1e3:.1f} ms per solve", flush=True)
            if "s_per_solve" not in best or s < best["s_per_solve"]:
                best["cont_stride"], best["s_per_solve"] = st, s
        del B
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    if splits:
        tuning = (best["tile_rows"], best["unroll"], best["unroll"], 0)
        if "s_per_solve" not in best:
            B = MeshBatch(meshes, which, mode=mode, nsp=best["nsp"], method=method, tuning=tuning)
            best["s_per_solve"] = _time_solve(B, tol)
            table.append({"nsp": best["nsp"], "tile_rows": best["tile_rows"], "unroll": best["unroll"],
                          "s_per_solve": best["s_per_solve"]})
            del B
        cols = _ncols(meshes, which)
        for wide, thr in splits:
            parts = SplitBatch.width_parts(cols, wide=wide, narrow=best["nsp"], threshold=thr)
            if len(parts) < 2 and parts[0][1] == best["nsp"]:
                continue                      # the single handle itself
            S = SplitBatch(meshes, which, mode=mode, parts=parts, method=method, tuning=tuning)
            if "cont_stride" in best:
                S.set_cont_stride(best["cont_stride"])
            s = _time_solve(S, tol)
            table.append({"split": [wide, thr], "parts": [[len(i), w] for i, w in parts], "s_per_solve": s})
            if verbose:
                print(f"[tune] split {wide}-wide above {thr} columns: {s * 1e3:.1f} ms per solve", flush=True)
            if s < best["s_per_solve"]:
                best["split"], best["s_per_solve"] = [wide, thr], s
            del S
            gc.collect()
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
    if use_cache:
        cache[key] = {"best": best, "table": table}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as fh:
                json.dump(cache, fh)
        except OSError:
            pass
    return best, table


def tuned_batch(meshes, which="s", mode="mixed", method="cocg", **kw):
    """A MeshBatch (or, when a partition was tuned faster, a SplitBatch) with
    the tuned group width, kernel shapes and (when tuned) continuation
    spacing."""
    meshes = list(meshes)
    best, _ = tune(meshes, which, mode, method, **kw)
    tuning = (best["tile_rows"], best["unroll"], best["unroll"], 0)
    if "split" in best:
        wide, thr = best["split"]
        parts = SplitBatch.width_parts(_ncols(meshes, which), wide=wide, narrow=best["nsp"], threshold=thr)
        B = SplitBatch(meshes, which, mode=mode, parts=parts, method=method, tuning=tuning)
    else:
        B = MeshBatch(meshes, which, mode=mode, nsp=best["nsp"], method=method, tuning=tuning)
    if "cont_stride" in best:
        B.set_cont_stride(best["cont_stride"])
    return B
