"""Benchmark: batched multi-frequency forward solve at 269k nodes on one MI355X.

Workload: 80x80x40 mm slab, 1 mm Kuhn grid (269,001 nodes, 1,536,000 tets),
mua 0.01/mm (6 mm sphere of 0.03/mm), musp 1.0/mm, n 1.4; 32 sources, 64
frequencies 0-3.15 GHz in 50 MHz steps = 2048 complex right-hand sides.

One step = FSAI factor setup for all 64 frequencies + the batched complex64
FSAI-preconditioned COCG inside FP64 iterative refinement to a relative
residual of 5e-8, with frequency continuation (each later-level frequency
started from the interpolation of neighbours solved before it), from
device-resident A and Q to converged fluence, + fc2td of the nodal fluence
to 256 time bins on FP64 MFMA.  The timed steps run the production path (the
captured Krylov iteration replayed as a HIP graph, no per-launch events).
value = 2048 * steps / timed seconds (RHS solves/s).  5e-8 is the loosest
residual whose fluence error against the FP64 solve (largest over the 2048
RHS, measured in the same run) stays a factor 3 below the 1e-6 target.

Also measured in the same run (reported beside the metric):
  * a separate profiled step (HIP events around every launch) for the
    per-class kernel times and the roofline of the dominant class,
  * the device copy rate and the fc2td DFT alone, with its own error (FP64
    fc2td of the same field) and the end-to-end TD error against the FP64
    field,
  * the mixed solve with BiCGSTAB (the reference's method) and without
    frequency continuation,
  * the FP64 fused solve at 1e-12 (the fluence error reference) and at the
    mixed tolerance,
  * the reference-structure FP64 solve (separate passes, lockstep, host
    checks, x0 = 0) at 1e-12 (the ">= 20x" denominator) and at the mixed
    tolerance (the equal-tolerance ratio),
  * the adjoint Jacobian: forward+adjoint (4096 RHS) mixed solve plus the
    fused J-TG kernel (1024 pairs x 8 gates x 269,001 nodes), and the voxel
    basis (mesh2grid at 1 and 2 mm, J-TG over the 2 mm voxels),
  * forward + adjoint + J-TG in the mixed form (two tolerance settings)
    against the reference-structure FP64 form, with the mixed J's error
    against the FP64 pipeline's J at sampled nodes,
  * the 48-mesh uneven batch (autotuned shapes, FP64 comparison),
  * the exact-derivative J_TG^T ybar against central differences at 50
    nodes of the 269k mesh (config 3),
  * the weighted LM reconstruction of two inclusions (config 5),
  * the oracle's C++/OpenMP FP64 FSAI-BiCGSTAB on every host core for one
    frequency x 32 sources (the CPU baseline).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0            # MI355X HBM3E
FP64_MFMA_PEAK_TFLOPS = 78.6     # MI355X dense FP64 matrix


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(small=False):
    from dogpup_amd import meshgen, mesh as dm
    t0 = time.time()
    if small:
        m, src, det, link = meshgen.slab_60x60x30_h2()
        nf = 8
    else:
        m, src, det, link = meshgen.slab_80x80x40_h1(seed=0)
        # nodes on the solver's Morton curve (iso2mesh's sortmesh does the
        # same for locality): the solves then skip their row permutations
        m = meshgen.sorted_mesh(m)
        nf = 64
    t = np.arange(0, 1001) * 20e-12                     # T = 20 ns -> 50 MHz spacing
    tpsf = meshgen.gaussian_tpsf(t, 250e-12)
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
    opt = dm.DOptode(src, det, link, 1e3, tpsf, t, nf)
    D.add_optode(opt)
    D.dxyz = m.h                                         # voxel side of the nodal basis
    ident = bool(np.array_equal(D.solver_order, np.arange(D.nn)))
    log(f"mesh {D.nn} nodes, {D.elem.shape[0]} tets, nnz(A) {D.nnzA}, nnz(G) {D.nnzG}, "
        f"host prep {time.time() - t0:.1f} s, node order = solver order: {ident}")
    return D, opt, t


def timed_solves(D, which, mode, tol, steps, warmup, profile=False, method=None, post=None, cont=None):
    """Run warmup + timed steps of (FSAI setup + solve [+ post(field)]);
    returns (seconds per step, field of the last step, per-class profile or
    None)."""
    import torch
    from dogpup_amd import _lib
    ncols = sum(len((D.optode.s_bary if w == "s" else D.optode.d_bary)[0]) for w in which)
    S = D.solver(ncols, mode, method)
    phi = None
    if warmup == 0:
        # a cheap untimed solve (two iterations) so that the timed one finds
        # the handle's buffers touched and its graphs captured
        D.solve_point_sources(which, mode=mode, tol=0.5, maxit=2, method=method, cont=cont)
    for _ in range(warmup):
        S._version = None
        phi = D.solve_point_sources(which, mode=mode, tol=tol, method=method, cont=cont)
        if post:
            post(phi)
    if profile:
        _lib.call("dogpup_solver_set_profile", S.h, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        S._version = None                       # FSAI setup is part of every step
        phi = D.solve_point_sources(which, mode=mode, tol=tol, method=method, cont=cont)
        if post:
            post(phi)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    prof = None
    if profile:
        p = np.zeros(12)
        _lib.call("dogpup_solver_profile", S.h, _lib.ptr(p), 12)
        _lib.call("dogpup_solver_set_profile", S.h, 0)
        prof = p.reshape(4, 3)
    return dt, phi, prof


def copy_bandwidth(nbytes=2 << 30, reps=5):
    """Device-to-device copy rate of a plain copy of nbytes (read + write
    counted), the practical HBM ceiling the kernels are compared with."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3) / 1e9


def release(D):
    import torch
    D._solvers.clear()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def rel_err_per_rhs(a, b):
    import torch
    num = torch.linalg.vector_norm(a - b, dim=1)
    den = torch.linalg.vector_norm(b, dim=1)
    return (num / den.clamp_min(1e-300))


def cpu_baseline(D, k=32, tol=1e-12):
    """The oracle's OpenMP FP64 restatement of the reference path for one
    frequency (index k, 1.6 GHz by default: the middle of the band) and all
    32 sources at 269,001 nodes, on every host core this process may use:
    FSAI setup (FSAIP_CUDA.cu:28-160) and reference-structure BiCGSTAB
    (dMesh.m:1283-1341) to the reference's 1e-12.  RHS/s = 32 / (setup +
    solve); the 2048-RHS job is 64 such frequencies."""
    import scipy.sparse as sp
    from oracle import cpu as ocpu, optode as oopt
    cores = len(os.sched_getaffinity(0))
    rA, cA, K, M = D.system_csr()
    n = D.nn
    Kc = sp.csr_matrix((K.cpu().numpy(), cA.cpu().numpy(), rA.cpu().numpy()), shape=(n, n))
    Mc = sp.csr_matrix((M.cpu().numpy(), cA.cpu().numpy(), rA.cpu().numpy()), shape=(n, n))
    opt = D.optode
    Ak = (Kc + 1j * (k * opt.df / D.c) * Mc).tocsr()
    rG = D._dev["rG"].cpu().numpy()
    cG = D._dev["cG"].cpu().numpy()
    sid, sw = opt.s_bary
    ns = len(sid)
    b = np.zeros((n, ns), complex)
    for s in range(ns):
        b[D.elem[sid[s]], s] = sw[s] * opt.s_fpsf[k]
    t0 = time.perf_counter()
    g = ocpu.fsai_setup(Ak, rG, cG, nthreads=cores)
    t_fsai = time.perf_counter() - t0
    t0 = time.perf_counter()
    x, it, rr = ocpu.bicgstab(Ak, rG, cG, g, b, tol=tol, maxit=1000, nthreads=cores)
    t_solve = time.perf_counter() - t0
    return {"value": ns / (t_fsai + t_solve), "unit": "RHS solves/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 FSAI setup + reference-structure BiCGSTAB (C++/OpenMP, {cores} threads), "
                      f"1 frequency ({k * 50} MHz) x {ns} sources at {n:,} nodes, tol {tol:g}: {it} iterations, "
                      f"{t_fsai:.2f} s setup + {t_solve:.1f} s solve, final relres {rr:.1e}"}


def batch_leg(tol):
    """The 48-mesh uneven batch: host prep, autotuned shapes (group width,
    tile, continuation spacing, and a split into parts of their own group
    width), one timed mixed solve (FSAI setup + solve) of all 38,431 RHS, the
    same with the library's default shapes in one handle, and the
    mixed-vs-FP64 error of every instance."""
    import gc
    import torch
    from dogpup_amd import autotune, batch, meshgen, mesh as dm
    t0 = time.time()
    specs = meshgen.batch_48_uneven()
    t = np.arange(0, 1001) * 20e-12
    tpsf = meshgen.gaussian_tpsf(t, 250e-12)
    meshes = []
    for sp in specs:
        m = meshgen.sorted_mesh(sp["mesh"])       # node order = solver order, as for the 269k-node mesh
        D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
        D.add_optode(dm.DOptode(sp["src"], sp["det"], sp["link"], 1e3, tpsf, t, sp["nf"]))
        meshes.append(D)
    nrhs = sum(len(sp["src"]) * sp["nf"] for sp in specs)
    out = {"instances": len(meshes), "nodes": int(sum(D.nn for D in meshes)), "rhs": nrhs,
           "node_order_is_solver_order": all(np.array_equal(D.solver_order, np.arange(D.nn)) for D in meshes),
           "node_rhs": float(sum(D.nn * len(sp["src"]) * sp["nf"] for D, sp in zip(meshes, specs))),
           "host_prep_s": time.time() - t0}
    log(f"batch: {len(meshes)} meshes, {out['nodes']} nodes, {nrhs} RHS; autotuning")
    t1 = time.time()
    best, table = autotune.tune(meshes, "s", "mixed", "cocg", use_cache=False, widths=(8, 16), tiles=(512, 1024),
                                unrolls=(8,), iters=4, strides=(8, 16, 32), splits=((32, 24), (32, 16)), tol=tol)
    out["tune_s"] = time.time() - t1
    out["tuned"] = best
    out["tune_table_ms_per_iter"] = {f"nsp{r['nsp']}_tile{r['tile_rows']}_u{r['unroll']}": r["s_per_iter"] * 1e3
                                     for r in table if "s_per_iter" in r}
    out["tune_table_ms_per_solve"] = {f"cont_stride{r['cont_stride']}": r["s_per_solve"] * 1e3
                                      for r in table if "cont_stride" in r}
    out["tune_table_ms_per_solve"].update({f"split_{r['split'][0]}wide_above{r['split'][1]}": r["s_per_solve"] * 1e3
                                           for r in table if "split" in r})

    def timed(B):
        b = B.rhs()
        B.reset_matrix()
        B.solve(b, tol=tol)                      # warm-up (graph capture)
        torch.cuda.synchronize()
        B.reset_matrix()                         # FSAI setup is part of the timed solve
        t2 = time.perf_counter()
        f = B.solve(b, tol=tol)
        torch.cuda.synchronize()
        return time.perf_counter() - t2, f, b

    default_nsp = batch.MeshBatch.default_width(meshes, [len(sp["src"]) for sp in specs],
                                                [sp["nf"] for sp in specs], "mixed")
    log(f"batch: library default shapes (nsp {default_nsp})")
    B = batch.MeshBatch(meshes, "s", "mixed", nsp=default_nsp, method="cocg")
    dt_def, _, b = timed(B)
    del B, b, _
    gc.collect()
    torch.cuda.empty_cache()
    log(f"batch: tuned shapes {best}")
    tuning = (best["tile_rows"], best["unroll"], best["unroll"], 0)
    if "split" in best:
        parts = batch.SplitBatch.width_parts([len(sp["src"]) for sp in specs], wide=best["split"][0],
                                             narrow=best["nsp"], threshold=best["split"][1])
        B = batch.SplitBatch(meshes, "s", "mixed", parts=parts, method="cocg", tuning=tuning)
        out["tuned_parts"] = [[len(i), w] for i, w in parts]
    else:
        B = batch.MeshBatch(meshes, "s", "mixed", nsp=best["nsp"], method="cocg", tuning=tuning)
    B.set_cont_stride(best.get("cont_stride", 0))
    dt, fields, b = timed(B)
    out["mixed_s"] = dt
    out["rhs_per_s"] = nrhs / dt
    out["default_shapes_s"] = dt_def
    out["tuning_speedup"] = dt_def / dt
    out["max_iters"] = int(max(f.iters.max() for f in fields))
    out["status"] = sorted({int(v) for f in fields for v in np.unique(f.status)})
    xm = [f.to_reference_layout().clone() for f in fields]     # the part widths may differ from the FP64 run's
    del B, fields, b
    gc.collect()
    torch.cuda.empty_cache()
    log("batch: FP64 reference solve")
    B64 = batch.MeshBatch(meshes, "s", "fp64", nsp=best["nsp"], method="cocg")
    b = B64.rhs()
    t3 = time.perf_counter()
    f64 = B64.solve(b, tol=1e-12)
    torch.cuda.synchronize()
    out["fp64_s"] = time.perf_counter() - t3
    errs = []
    for a, f in zip(xm, f64):                  # one mesh's reference-layout copy at a time
        r = f.to_reference_layout()
        errs.append(float(torch.linalg.vector_norm(a - r) / torch.linalg.vector_norm(r)))
        del r
    out["mixed_vs_fp64_rel_err_per_instance"] = errs
    out["mixed_vs_fp64_rel_err_max"] = max(errs)
    del B64, f64, xm, b
    gc.collect()
    torch.cuda.empty_cache()
    return out


def td_error(td_sample, vs, phi64, opt, t_td):
    """TD curves of the timed step (FP64-MFMA fc2td of the mixed field, FP32
    output) at the nodes vs against the FP64 fc2td (dogpup_fc2td) of the
    field phi64 (the mixed field itself: the transform's own error; the FP64
    field: the end-to-end error): Frobenius relative error, and per curve
    (node x source) the largest error relative to that curve's peak (max and
    median), also over the curves whose peak is at least 1e-2 of their
    source's largest curve peak."""
    import torch
    from dogpup_amd import transforms as tr
    tm = td_sample.double()                                                    # V x NT x NS
    t64 = tr.fc2td(phi64.to_reference_layout()[vs], opt.fAxis, t_td, 1)        # V x NT x NS
    d = (tm - t64).abs()
    pk = t64.abs().amax(1)                                                     # V x NS
    curve = d.amax(1) / pk
    strong = pk >= 1e-2 * pk.amax(0, keepdim=True)
    return {"sampled_nodes": int(vs.numel()), "rel_err_fro": float(torch.linalg.vector_norm(tm - t64)
                                                          / torch.linalg.vector_norm(t64)),
            "curve_err_over_peak_max": float(curve.max()), "curve_err_over_peak_median": float(curve.median()),
            "curve_err_over_peak_max_strong_curves": float(curve[strong].max()),
            "strong_curves": int(strong.sum()), "curves": int(curve.numel()),
            "err_over_global_max": float(d.max() / t64.abs().max())}


def target_leg(D, gfc, tol, method, adj_tol, nsample=4000, seed=1):
    """Forward + adjoint solve of every source and detector at every
    frequency (4096 RHS) and J-TG (1024 pairs x 8 gates x 269k nodes): the
    mixed form (COCG with frequency continuation, warm handle, FSAI setup
    inside the timing, as in the headline step) against the
    reference-structure FP64 form (BiCGSTAB from x0 = 0 to 1e-12, the
    tolerance dMesh.m:1243 hard-codes; same J-TG kernel).  The mixed form
    runs at three settings: the source columns to tol and the detector
    (adjoint) columns to adj_tol or to adj_tol / 10, and every column to
    1e-8.  `speedup` is the fastest setting whose J meets the 1e-4 target
    on every gate (largest error relative to the gate's largest entry).
    The mixed J is compared with J of the FP64 field,
    J64[v, g, m] = Re sum_k C_gk phi(v, k, s_m) psi(v, k, d_m) in complex128
    at nsample seeded nodes, over every (pair, gate) column: the Frobenius
    relative error, and per gate the largest error of a column relative to
    that column's largest entry and to the gate's largest entry."""
    import torch
    from dogpup_amd.jacobian import jacobian_tg, tg_coefficients
    log("forward + Jacobian: warm mixed runs")
    D.solve_point_sources("sd", mode="mixed", tol=tol, method=method)      # handle, graph capture
    S = D.solver(sum(len(b[0]) for b in (D.optode.s_bary, D.optode.d_bary)), "mixed", method)
    rng = np.random.default_rng(seed)
    vs = torch.as_tensor(np.sort(rng.choice(D.nn, nsample, replace=False)), device="cuda")

    def run_mixed(ftol, atol):
        S._version = None                                                   # FSAI setup inside the timing
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = D.solve_point_sources("sd", mode="mixed", tol=ftol, method=method, adj_tol=atol)
        J, info = jacobian_tg(D, gfc, field=f)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, f, J, info
    ns = len(D.optode.s_bary[0])
    settings = {"tol_adj": (tol, adj_tol), "tol_adj10": (tol, adj_tol / 10), "uniform_1e-8": (1e-8, None)}
    mixed, Js = {}, {}
    for key, (ft, at) in settings.items():
        run_mixed(ft, at)                                                   # warm-up at this setting
        dt, fm, Jm, im = run_mixed(ft, at)
        mixed[key] = {"s": dt, "fwd_tol": ft, "adj_tol": at if at is not None else ft,
                      "max_iters": int(fm.iters.max()), "mean_iters_forward": float(fm.iters[:, :ns].mean()),
                      "mean_iters_adjoint": float(fm.iters[:, ns:].mean()), "jtg_kernel_s": im["kernel_s"]}
        Js[key] = Jm[vs].double()
        del fm, Jm
    release(D)
    log("forward + Jacobian: reference-structure FP64 run")
    D.solve_point_sources("sd", mode="plain", tol=0.5, maxit=2)          # untimed: handle buffers touched
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fp = D.solve_point_sources("sd", mode="plain", tol=1e-12)
    Jp, _ = jacobian_tg(D, gfc, field=fp)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    its_p = int(fp.iters.max())
    opt = D.optode
    R = fp.to_reference_layout()[vs]                                   # V x NF x NCOLS
    ls = torch.as_tensor(opt.link[:, 0].astype(np.int64), device="cuda")
    ld = torch.as_tensor(opt.link[:, 1].astype(np.int64) + ns, device="cuda")
    coef = tg_coefficients(opt, gfc, D.dxyz)
    J64 = torch.einsum("gk,vkm->vgm", coef, R[:, :, ls] * R[:, :, ld]).real
    Jp_s = Jp[vs].double()
    del Jp, fp, R

    def errs(J):
        d = (J - J64).abs()
        col = (d.amax(0) / J64.abs().amax(0).clamp_min(1e-300)).amax(1)      # per gate
        gate = d.amax(0).amax(1) / J64.abs().amax(0).amax(1)
        return (float(torch.linalg.vector_norm(J - J64) / torch.linalg.vector_norm(J64)),
                col.tolist(), gate.tolist())
    for key in settings:
        fro, col, gate = errs(Js[key])
        mixed[key].update(speedup=t_plain / mixed[key]["s"], J_rel_err_fro=fro, J_col_rel_err_max_per_gate=col,
                          J_err_over_gate_max_per_gate=gate, meets_J_target=bool(max(gate) <= 1e-4))
    fro_p, col_p, gate_p = errs(Jp_s)
    ok = [k for k in settings if mixed[k]["meets_J_target"]]
    best = min(ok, key=lambda k: mixed[k]["s"]) if ok else None
    out = {"rhs": 2 * 2048, "plain_s": t_plain, "plain_iters": its_p, "mixed": mixed,
           "speedup": mixed[best]["speedup"] if best else None, "speedup_setting": best,
           "J_target": "largest |J - J64| of every gate <= 1e-4 of the gate's largest |J64|",
           "speedup_uniform_1e-8": mixed["uniform_1e-8"]["speedup"],
           "sampled_nodes": nsample,
           "fp64_field_J_rel_err_fro": fro_p, "fp64_field_J_col_rel_err_max_per_gate": col_p,
           "note": "mixed = COCG c64 + FP64 refinement with frequency continuation, warm handle with FSAI setup; "
                   "tol_adj: sources to tol, detectors to adj_tol; plain = reference-structure FP64 BiCGSTAB to "
                   "1e-12 from x0 = 0; J64 = complex128 J-TG of the plain field at sampled nodes (fp64_field_J_* "
                   "is the J-TG kernel's own error on that field)"}
    return out


def fd_leg(n_nodes=50, seed=0, delta=1e-5):
    """Config 3 derivative check at full size: on the 269,001-node mesh, the
    exact-derivative vector-Jacobian product g = J_TG^T ybar (kappa fixed,
    dogpup_jacobian_vjp) against central differences of ybar . y(mua) at 50
    seeded nodes under the optodes, and the reference's nodal-product J at the
    same nodes.  A reduced optode set (4 sources x 4 detectors, 8
    frequencies, 8 gates) keeps the 100 FP64 forward solves to about a
    minute."""
    import torch
    from dogpup_amd import jacobian as jac, meshgen, mesh as dm
    from dogpup_amd.transforms import td2fc
    t = np.arange(0, 1001) * 20e-12
    m, src, det, _, = meshgen.slab_80x80x40_h1(seed=0)
    m = meshgen.sorted_mesh(m)                    # node order = solver order, as in the other legs
    src = src[[0, 9, 18, 27]]
    det = det[[4, 13, 22, 31]]
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
    D.add_optode(dm.DOptode(src, det, meshgen.all_pairs(4, 4), 1e3, meshgen.gaussian_tpsf(t, 250e-12), t, 8))
    gates = meshgen.boxcar_gates(t, 8, 500e-12, 200e-12, start=500e-12)
    gfc = td2fc(torch.as_tensor(gates), D.optode.fAxis, t, 1)
    rng = np.random.default_rng(seed)
    ybar = rng.normal(size=(16, 8))
    kappa0 = D.kappa.copy()
    mua0 = D.mua.copy()
    t0 = time.perf_counter()
    f = D.solve_point_sources("sd", mode="fp64", tol=1e-13)
    g_exact = jac.vjp(D, f, gfc, torch.as_tensor(ybar, device="cuda"), exact=True).cpu().numpy()
    g_nodal = jac.vjp(D, f, gfc, torch.as_tensor(ybar, device="cuda"), exact=False, dxyz=1.0).cpu().numpy()
    del f
    lo = np.array([15.0, 15.0, 40.0 - 15.0])
    hi = np.array([65.0, 65.0, 40.0 - 2.0])
    cand = np.flatnonzero(np.all((D.node >= lo) & (D.node <= hi), axis=1))
    nodes = np.sort(rng.choice(cand, n_nodes, replace=False))

    def objective(mu):
        D.update_properties(np.stack([mu, D.musp], 1), kappa=kappa0)
        fs = D.solve_point_sources("s", mode="fp64", tol=1e-13)
        return float((jac.tg_data(D, fs, gfc).cpu().numpy() * ybar).sum())

    fd = np.zeros(n_nodes)
    for q, v in enumerate(nodes):
        vals = []
        for sgn in (1.0, -1.0):
            mu = mua0.copy()
            mu[v] += sgn * delta
            vals.append(objective(mu))
        fd[q] = (vals[0] - vals[1]) / (2.0 * delta)
    D.update_properties(np.stack([mua0, D.musp], 1), kappa=kappa0)
    dt = time.perf_counter() - t0
    scale = np.abs(g_exact[nodes]).max()
    err = np.abs(fd - g_exact[nodes])
    rel_nodal = np.abs(g_nodal[nodes] - g_exact[nodes]) / np.maximum(np.abs(g_exact[nodes]), 1e-300)
    return {"nodes": int(n_nodes), "mesh_nodes": int(D.nn), "delta": delta, "seconds": dt,
            "max_abs_err_over_max_g": float(err.max() / scale),
            "median_rel_err": float(np.median(err / np.maximum(np.abs(g_exact[nodes]), 1e-300))),
            "nodal_product_vs_exact_median_rel": float(np.median(rel_nodal)),
            "nodal_product_vs_exact_max_rel": float(rel_nodal.max())}


def shard_leg(D, tol, method, x_ref, ngpu=8):
    """One GPU's share of the job split over `ngpu` GPUs, in three layouts of
    ngpu shards of NS x NF / ngpu right-hand sides each: contiguous frequency
    blocks x every source; source blocks x every frequency; and source blocks
    x frequency halves.  Each shard has its own handle (own G, own
    continuation levels over its frequencies: a shard needs nothing from
    another during its solve).  Every shard is run here on this one GPU, one
    after the other: a warm-up solve (graph capture), then one timed step of
    FSAI setup + mixed solve, the headline step without fc2td (which in the
    split job runs after an exchange by node range).  Each shard's field is
    compared with the full solve's columns."""
    import copy
    import torch
    opt = D.optode
    nf, ns = opt.Nf, len(opt.s_bary[0])
    fall, sall = np.arange(nf), np.arange(ns)
    layouts = {
        "frequency_blocks": [(sall, fall[r * nf // ngpu:(r + 1) * nf // ngpu]) for r in range(ngpu)],
        "source_blocks": [(sall[r * ns // ngpu:(r + 1) * ns // ngpu], fall) for r in range(ngpu)],
        "source_blocks_x_frequency_halves": [(sall[q * 2 * ns // ngpu:(q + 1) * 2 * ns // ngpu],
                                              fall[h * nf // 2:(h + 1) * nf // 2])
                                             for q in range(ngpu // 2) for h in range(2)],
    }
    out = {}
    try:
        for name, shards in layouts.items():
            res = []
            for sidx, freqs in shards:
                sub = copy.copy(opt)
                sub.s_bary = (opt.s_bary[0][sidx], opt.s_bary[1][sidx])
                D.optode = sub
                fq = None if len(freqs) == nf else freqs
                D.solve_point_sources("s", mode="mixed", tol=tol, method=method, freqs=fq)
                S = D.solver(len(sidx), "mixed", method, fq)
                S._version = None                       # FSAI setup is part of the step
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                fw = D.solve_point_sources("s", mode="mixed", tol=tol, method=method, freqs=fq)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                assert fw.data.shape[0] == len(freqs) and fw.nsp >= len(sidx), "one column group per frequency"
                ref = x_ref[torch.as_tensor(freqs, device=x_ref.device)][:, :, int(sidx[0]):int(sidx[-1]) + 1]
                err = rel_err_per_rhs(fw.data[:, :, :len(sidx)].permute(0, 2, 1).reshape(-1, D.nn),
                                      ref.permute(0, 2, 1).reshape(-1, D.nn))
                res.append({"sources": [int(sidx[0]), int(sidx[-1])], "freqs": [int(freqs[0]), int(freqs[-1])],
                            "group_width": int(fw.nsp), "s": dt, "max_iters": int(fw.iters.max()),
                            "total_iters": int(fw.iters.sum()), "vs_full_solve_rel_diff_max": float(err.max())})
                del fw
                D.optode = opt
                release(D)
            ts = [r["s"] for r in res]
            out[name] = {"shards": res, "max_s": max(ts), "sum_s": sum(ts)}
    finally:
        D.optode = opt
    out["rhs_per_shard"] = ns * nf // ngpu
    out["note"] = (f"one GPU's share of the 2048-RHS job split over {ngpu} GPUs, in three layouts; every shard "
                   f"timed on this one GPU (step = FSAI setup + mixed solve, no fc2td); no multi-GPU run is made")
    return out


def lm_leg(max_iter=10):
    """Config 5: weighted LM absorption reconstruction.  Data: FP64 solve on
    a 0.75 mm mesh of the slab with two inclusions, 32 x 32 pairs x 8 gates,
    Poisson noise at 1e6 peak counts; inversion on the 1 mm (269k-node) mesh
    from the homogeneous background, in the mixed COCG form with frequency
    continuation at three settings (every column to 1e-8; forward columns
    to 5e-8 and adjoint columns to 1e-6 or 1e-7) and in the reference-structure FP64
    form (1e-12); the reconstructed mua of each mixed setting against the
    FP64 form's."""
    import gc
    import torch
    from dogpup_amd import lm, meshgen, mesh as dm
    t = np.arange(0, 1001) * 20e-12
    tpsf = meshgen.gaussian_tpsf(t, 250e-12)
    gates = meshgen.boxcar_gates(t, 8, 500e-12, 200e-12, start=500e-12)
    out = {}
    t0 = time.time()
    mf, src, det, link, inc = meshgen.slab_80x80x40_two_inclusions(0.75, seed=0)
    mf = meshgen.sorted_mesh(mf)
    Df = dm.DMesh(mf.node, mf.elem, mf.nr, np.stack([mf.mua, mf.musp], 1))
    Df.add_optode(dm.DOptode(src, det, link, 1e3, tpsf, t, 64))
    log(f"LM data mesh {Df.nn} nodes (host prep {time.time() - t0:.1f} s); FP64 forward solve")
    t1 = time.time()
    data = lm.simulate_tg_data(Df, gates, mode="fp64", tol=1e-10, peak_counts=1e6, seed=0)
    out["data_mesh_nodes"] = Df.nn
    out["data_s"] = time.time() - t1
    del Df
    gc.collect()
    torch.cuda.empty_cache()
    m, src, det, link, inc = meshgen.slab_80x80x40_two_inclusions(1.0, seed=0)
    m = meshgen.sorted_mesh(m)                    # node order = solver order (no row permutations)
    truth = m.mua.copy()
    xs = np.arange(8.0, 73.0, 8.0)
    pos_s = np.stack(np.meshgrid(xs, xs, [30.0, 25.0, 20.0], indexing="ij"), -1).reshape(-1, 3)
    runs = (("mixed", "cocg", 1e-8, None), ("mixed_tol_adj", "cocg", 5e-8, 1e-6), ("mixed_tol_adj10", "cocg", 5e-8, 1e-7),
            ("plain", "bicgstab", 1e-12, None))
    for mode, method, tol, atol in runs:
        D = dm.DMesh(m.node, m.elem, m.nr, np.stack([np.full(m.node.shape[0], 0.01), m.musp], 1))
        D.add_optode(dm.DOptode(src, det, link, 1e3, tpsf, t, 64))
        D.dxyz = 1.0
        log(f"LM {mode} ({method}), {max_iter} iterations at most")
        t2 = time.time()
        # lam = (LM damping, spot reshaping); 1e-1 keeps reshape_J's inverse of the
        # 8192-row J J^T stable under the Poisson noise (1e-2 diverged)
        r = lm.wghtd_lm_recon(D, data, gates, pos_s, (1e2, 1e-1), mode=mode.split("_")[0], tol=tol,
                              max_iter=max_iter, method=method, adj_tol=atol)
        wall = time.time() - t2
        inside = [np.linalg.norm(m.node - np.asarray(c)[None, :], axis=1) <= rad for c, rad, _ in inc]
        out[mode] = {"s": wall, "tol": tol, "adj_tol": atol, "iterations": r.count, "exit": r.exit,
                     "errPerc": r.errPerc,
                     "rmse": float(np.sqrt(np.mean((r.mua - truth) ** 2))),
                     "contrast": [float(r.mua[k].mean() / 0.01) for k in inside],
                     "true_contrast": [float(mu / 0.01) for _, _, mu in inc],
                     "t_forward_s": r.t_forward, "t_algebra_s": r.t_algebra}
        out[mode + "_mua"] = r.mua
        del D
        gc.collect()
        torch.cuda.empty_cache()
    b = out.pop("plain_mua")
    for key in ("mixed", "mixed_tol_adj", "mixed_tol_adj10"):
        a = out.pop(key + "_mua")
        out[key]["mua_rel_diff_vs_plain"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        out[key]["speedup_vs_plain"] = out["plain"]["s"] / out[key]["s"]
        out[key]["within_0.1pct_of_plain"] = bool(out[key]["mua_rel_diff_vs_plain"] <= 1e-3)
    out["mixed_vs_plain_mua_rel_diff"] = out["mixed"]["mua_rel_diff_vs_plain"]
    out["speedup_mixed_vs_plain"] = out["mixed"]["speedup_vs_plain"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="60x60x30 mm, 2 mm grid, 8 frequencies (quick check)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-jacobian", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 48-mesh uneven batch")
    ap.add_argument("--no-lm", action="store_true", help="skip the LM reconstruction (config 5, about a minute)")
    ap.add_argument("--no-fd", action="store_true",
                    help="skip the exact-derivative J_TG^T ybar check against central differences at 50 nodes of "
                         "the 269k mesh (config 3, about half a minute)")
    ap.add_argument("--no-profile", action="store_true", help="skip the separate profiled step (per-class times)")
    ap.add_argument("--no-shard", action="store_true",
                    help="skip the shards (one GPU's share of the job split over 8 GPUs: by frequency, by source, by both)")
    ap.add_argument("--method", default="cocg", choices=["bicgstab", "cocg"],
                    help="Krylov method of the headline mixed solve")
    ap.add_argument("--no-alt", action="store_true",
                    help="skip the mixed solves with the other method, with Lagrange continuation and "
                         "without continuation")
    ap.add_argument("--adj-tol", type=float, default=1e-6,
                    help="relative FP64 residual of the adjoint columns in the forward + Jacobian leg "
                         "(the Jacobian's 1e-4 error target)")
    ap.add_argument("--tol", type=float, default=5e-8,
                    help="relative FP64 residual of the mixed solve (chosen to meet 1e-6 fluence error)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; multi-GPU sharding is design-only "
                                   "(frequency-sharded solves need no communication)", "n_gpus": a.gpus}))
        return 2
    import torch
    assert torch.cuda.is_available(), "bench needs a GPU"
    copy_gbs = copy_bandwidth()
    log(f"device copy {copy_gbs:.0f} GB/s")
    D, opt, t = build_workload(a.small)
    ns = len(opt.s_bary[0])
    nrhs = ns * opt.Nf
    extra = {}

    # headline: the production path (hipGraph replay of the Krylov iteration,
    # no per-launch events), one step = FSAI setup + mixed solve + fc2td of
    # the nodal fluence to 256 bins of 20 ps
    from dogpup_amd import transforms as tr
    t_td = np.arange(256) * 20e-12
    td_out = {}

    def to_td(phi):
        td_out["td"] = tr.fc2td_field(phi, opt.fAxis, t_td)

    log("mixed solve + fc2td (timed, production graph path)")
    dt_mixed, phi_mixed, _ = timed_solves(D, "s", "mixed", a.tol, a.steps, a.warmup, profile=False,
                                          method=a.method, post=to_td)
    value = nrhs / dt_mixed
    extra["mixed_s_per_solve"] = dt_mixed
    extra["mixed_max_iters"] = int(phi_mixed.iters.max())
    extra["mixed_total_iters"] = int(phi_mixed.iters.sum())
    extra["mixed_iters_per_freq"] = phi_mixed.iters.max(axis=1).tolist()
    extra["mixed_status"] = np.unique(phi_mixed.status).tolist()
    extra["mixed_relres_max"] = float(phi_mixed.relres.max())
    extra["mixed_sweeps"] = phi_mixed.stats["sweeps"]
    x_mixed = phi_mixed.data.clone()
    prof = None
    if not a.no_profile:
        # per-class kernel times and the roofline from a separate step with
        # HIP events around every launch on the solver's stream
        log("profiled step (per-launch events)")
        dt_prof, _, prof = timed_solves(D, "s", "mixed", a.tol, 1, 0, profile=True, method=a.method, post=to_td)
        extra["profiled_step_s"] = dt_prof
    # the DFT alone (MFMA roofline): 2 * NT * 2NF flops and 16 NF + 4 NT bytes
    # per column
    ncol_td = phi_mixed.data.numel() // opt.Nf
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    e0.record()
    for _ in range(reps):
        tr.fc2td_field(phi_mixed, opt.fAxis, t_td)
    e1.record()
    torch.cuda.synchronize()
    ms_td = e0.elapsed_time(e1) / reps
    flops_td = 2.0 * t_td.size * 2 * opt.Nf * ncol_td
    bytes_td = ncol_td * (16.0 * opt.Nf + 4.0 * t_td.size)
    extra["fc2td"] = {"ms": ms_td, "TFLOPs": flops_td / ms_td / 1e9, "mfma_peak_TFLOPs": FP64_MFMA_PEAK_TFLOPS,
                      "mfma_frac": flops_td / ms_td / 1e9 / FP64_MFMA_PEAK_TFLOPS,
                      "GBps": bytes_td / ms_td / 1e6, "columns": ncol_td, "bins": int(t_td.size),
                      "note": "fc2td of the mixed-solve nodal fluence (32 sources, 269,001 nodes) to 256 bins "
                              "of 20 ps on FP64 MFMA (FP32 output); part of every timed step"}
    # the step's TD curves at seeded nodes, kept for the FP64 comparison, and
    # the DFT's own error: FP64 fc2td (dogpup_fc2td) of the same mixed field
    td_vs = torch.as_tensor(np.sort(np.random.default_rng(2).choice(D.nn, 4000, replace=False)), device="cuda")
    td_sample = tr.td_field_to_reference(td_out["td"], phi_mixed.ncols)[td_vs].clone()
    del td_out["td"]
    extra["td_transform_vs_fp64_same_field"] = td_error(td_sample, td_vs, phi_mixed, opt, t_td)
    release(D)
    if not a.no_alt:
        alt = "cocg" if a.method == "bicgstab" else "bicgstab"
        log(f"mixed solve, {alt}")
        dta, pha, _ = timed_solves(D, "s", "mixed", a.tol, 1, 1, method=alt)
        err = rel_err_per_rhs(pha.data.permute(0, 2, 1).reshape(-1, D.nn),
                              x_mixed.permute(0, 2, 1).reshape(-1, D.nn))
        extra[f"mixed_{alt}_s_per_solve"] = dta
        extra[f"mixed_{alt}_max_iters"] = int(pha.iters.max())
        extra[f"mixed_{alt}_iters_per_freq"] = pha.iters.max(axis=1).tolist()
        extra[f"mixed_{alt}_status"] = np.unique(pha.status).tolist()
        extra[f"mixed_{alt}_vs_{a.method}_rel_diff_max"] = float(err.max())
        del pha
        release(D)
        for cont, name in (("lagrange", "mixed_lagrange"), (False, "mixed_no_continuation")):
            log(f"mixed solve, continuation {cont}")
            dtn, phn, _ = timed_solves(D, "s", "mixed", a.tol, 1, 1, method=a.method, cont=cont)
            err = rel_err_per_rhs(phn.data.permute(0, 2, 1).reshape(-1, D.nn),
                                  x_mixed.permute(0, 2, 1).reshape(-1, D.nn))
            extra[f"{name}_s_per_solve"] = dtn
            extra[f"{name}_total_iters"] = int(phn.iters.sum())
            extra[f"{name}_iters_per_freq"] = phn.iters.max(axis=1).tolist()
            extra[f"{name}_vs_galerkin_rel_diff_max"] = float(err.max())
            del phn
            release(D)

    if not a.no_shard and not a.small:
        log("shards: one GPU's share of the job split over 8 GPUs, three layouts")
        extra["shards"] = shard_leg(D, a.tol, a.method, x_mixed)

    names = ["G/G^T SpMV", "A SpMV", "Krylov vector passes", "FP64 refinement passes"]
    achieved = byts = nl = ms = None
    cls = 0
    if prof is not None:
        cls = int(np.argmax(prof[:, 0]))
        ms, byts, nl = prof[cls]
        achieved = byts / (ms * 1e-3) / 1e9
        extra["kernel_classes"] = {names[c]: {"ms": prof[c, 0], "GB": prof[c, 1] / 1e9, "launches": int(prof[c, 2]),
                                              "GB/s": prof[c, 1] / max(prof[c, 0], 1e-9) / 1e6}
                                   for c in range(4) if prof[c, 2] > 0}
        extra["all_kernels_GBps"] = prof[:, 1].sum() / (prof[:, 0].sum() * 1e-3) / 1e9
    extra["copy_GBps"] = copy_gbs

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{a.method}_c64_slab269k_2048rhs.json")
    if prof is not None and not a.small and os.path.exists(tfile):
        with open(tfile) as fh:
            tcl = json.load(fh)["classes"].get(str(cls))
        if tcl:
            # traffic/algorithmic ratio of the class's full-activity launches (rocprofv3
            # PMC passes FETCH_SIZE and WRITE_SIZE, calibrated on a plain copy; made by
            # python -m dogpup_amd.prof_traffic), applied to this run's average launch
            traffic = {"bytes_per_launch": tcl["ratio"] * byts / max(nl, 1), "ratio": tcl["ratio"],
                       "source": os.path.relpath(tfile, ROOT)}
    if not a.no_fp64:
        log("FP64 fused solve")
        dt64, phi64, _ = timed_solves(D, "s", "fp64", 1e-12, 1, 0, post=to_td)
        td_out.pop("td", None)
        err = rel_err_per_rhs(x_mixed.reshape(x_mixed.shape[0], -1, x_mixed.shape[2]).permute(0, 2, 1).reshape(-1, D.nn),
                              phi64.data.permute(0, 2, 1).reshape(-1, D.nn))
        extra["fp64_fused_s_per_solve"] = dt64
        extra["fp64_fused_max_iters"] = int(phi64.iters.max())
        extra["fp64_iters_per_freq"] = phi64.iters.max(axis=1).tolist()
        extra["mixed_vs_fp64_rel_err_max"] = float(err.max())
        extra["mixed_vs_fp64_rel_err_median"] = float(err.median())
        extra["fluence_error_target"] = 1e-6
        extra["meets_fluence_error_target"] = bool(float(err.max()) <= 1e-6)
        extra["speedup_mixed_vs_fp64_fused"] = dt64 / dt_mixed
        extra["td_mixed_vs_fp64_field"] = td_error(td_sample, td_vs, phi64, opt, t_td)
        del phi64
        release(D)
        log(f"FP64 fused solve at the mixed form's tolerance {a.tol:g}")
        dt64e, phe, _ = timed_solves(D, "s", "fp64", a.tol, 1, 0, post=to_td)
        td_out.pop("td", None)
        extra["fp64_fused_equal_tol_s_per_solve"] = dt64e
        extra["speedup_mixed_vs_fp64_fused_equal_tol"] = dt64e / dt_mixed
        del phe
        release(D)
    if not a.no_plain:
        log("reference-structure FP64 solve")
        dtp, php, _ = timed_solves(D, "s", "plain", 1e-12, 1, 0, post=to_td)
        td_out.pop("td", None)
        extra["fp64_plain_s_per_solve"] = dtp
        extra["fp64_plain_iters"] = int(php.iters.max())
        extra["speedup_mixed_vs_fp64_plain"] = dtp / dt_mixed
        del php
        release(D)
        log(f"reference-structure FP64 solve at the mixed form's tolerance {a.tol:g}")
        dtpe, phpe, _ = timed_solves(D, "s", "plain", a.tol, 1, 0, post=to_td)
        td_out.pop("td", None)
        extra["fp64_plain_equal_tol_s_per_solve"] = dtpe
        extra["fp64_plain_equal_tol_iters"] = int(phpe.iters.max())
        extra["speedup_mixed_vs_fp64_plain_equal_tol"] = dtpe / dt_mixed
        if f"mixed_bicgstab_s_per_solve" in extra:
            extra["speedup_mixed_bicgstab_vs_fp64_plain"] = dtp / extra["mixed_bicgstab_s_per_solve"]
        del phpe
        release(D)
    if not a.no_jacobian:
        log("Jacobian: forward+adjoint mixed solve + fused J-TG")
        from dogpup_amd.jacobian import jacobian_tg, boxcar_gates_fc
        gfc = boxcar_gates_fc(opt, t, 8, 500e-12, 200e-12)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, info = jacobian_tg(D, gfc, mode="mixed", tol=a.tol, method=a.method, adj_tol=a.adj_tol)
        torch.cuda.synchronize()
        extra["jacobian_s"] = time.perf_counter() - t0
        extra["jacobian_solve_s"] = info["solve_s"]
        extra["jacobian_tg_kernel_s"] = info["kernel_s"]
        extra["jacobian_shape"] = list(J.shape)
        del J
        # voxel basis: mesh2grid on 1 mm and 2 mm grids over the slab (point
        # location of every voxel, Kuhn-split weights of every node), then
        # J-TG over the 2 mm grid's voxels from the same fields (m2g applied to
        # the whole forward+adjoint field first)
        log("voxel basis: mesh2grid 1 mm / 2 mm, J-TG on 2 mm voxels")
        ext = np.asarray(D.node.max(axis=0))
        for hv in (1.0, 2.0):
            reps = []
            for _ in range(3):              # best of three host-timed calls
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                D.mesh2grid(*(np.arange(0.0, e + 0.5 * hv, hv) for e in ext))
                torch.cuda.synchronize()
                reps.append(time.perf_counter() - t0)
            extra[f"mesh2grid_{int(hv)}mm_s"] = min(reps)
        extra["voxels_2mm"] = [int(D.basis.nv), int(D.basis.nv_in)]
        t0 = time.perf_counter()
        Jv, iv = jacobian_tg(D, gfc, field=info["field"])
        torch.cuda.synchronize()
        extra["jacobian_tg_voxel2mm_s"] = time.perf_counter() - t0
        extra["jacobian_voxel2mm_shape"] = list(Jv.shape)
        D.clear_grid()
        del Jv, iv, info
        release(D)
        if not a.no_plain and not a.small:
            extra["forward_plus_jacobian"] = target_leg(D, gfc, a.tol, a.method, a.adj_tol)
            release(D)
    if not a.no_batch and not a.small:
        release(D)
        log("48-mesh uneven batch")
        extra["batch48"] = batch_leg(a.tol)
    if not a.no_fd and not a.small:
        release(D)
        log("FD check of the Jacobian at 50 nodes (config 3)")
        extra["fd_check"] = fd_leg()
    if not a.no_lm and not a.small:
        release(D)
        log("LM reconstruction (config 5)")
        extra["lm"] = lm_leg()
    cpu = None
    if not a.no_cpu and not a.small:
        log("CPU baseline (oracle, C++/OpenMP on every host core)")
        try:
            cpu = cpu_baseline(D)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": str(e)}
    out = {
        "metric": "src x freq RHS solves/s at 269k nodes (mixed vs FP64, 1e-6 rel err); Jacobian s",
        "value": value, "unit": "RHS solves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": dt_mixed * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64+f64-refine", "data": "synthetic",
        "config": {"workload": "slab 80x80x40 mm, 1 mm Kuhn grid (269,001 nodes), 32 sources x 64 "
                               f"frequencies = 2048 RHS, FSAI-preconditioned {a.method.upper()} in complex64 "
                               f"with FP64 refinement to a {a.tol:g} FP64 relative residual and frequency "
                               f"continuation, then fc2td of the nodal fluence to 256 bins of 20 ps (FP64 MFMA)"
                   if not a.small else "slab 60x60x30 mm 2 mm grid, 16 sources x 8 frequencies",
                   "nodes": D.nn, "rhs": nrhs, "tol": a.tol, "method": a.method},
        "roofline": {"bound": "hbm", "kernel": names[cls], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_ratio": traffic["ratio"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "bytes_per_launch": byts / max(nl, 1) if prof is not None else None,
                     "avg_launch_ms": ms / max(nl, 1) if prof is not None else None,
                     "measured_in": "a separate profiled step (HIP events around every launch on the solver's "
                                    "stream); the timed steps replay the captured graph without events"},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out, default=float))
    return 0


if __name__ == "__main__":
    sys.exit(main())
