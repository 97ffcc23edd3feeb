"""Weighted Levenberg-Marquardt absorption reconstruction on the GPU path.

Host logic mirrors ``DOGPUP/reconstruction/wghtd_LM_recon.m:1-169`` (with
``id_thresh.m`` and ``dMesh.m:250-282`` target_spots).  The unknowns live on
the mesh's voxel basis when ``mesh.mesh2grid`` has attached one (the update
is solved for the in-mesh voxels, scaled by m2g*mua and mapped back with g2m,
:101-134), otherwise on the nodal basis (grid = mesh nodes: m2g = g2m = I,
gridinmesh all true).  Per iteration the heavy work runs in
libdogpup_hip.so:
  - one batched forward+adjoint solve (sources and detectors as columns);
  - the time-gated data fc2tg(meas_flu(phi));
  - the fused J-TG kernel with the Rytov scaling 1/data;
  - the normal matrices J J^T and Js J^T on FP64 MFMA (FP32 entries, exact
    products and FP64 accumulation);
  - the FP64 Cholesky solves of reshape_J and tik_inv (data-space form).
Row selection, logs, norms and the accept/reject logic are host glue, as in
the reference.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import recon
from .jacobian import jacobian_tg, tg_data
from .transforms import id_thresh, td2fc

DEV = "cuda"


def target_spots(node, pos, w, kind="gauss", inmesh=None):
    """Sensitivity target spots (dMesh.m:250-282) on the grid points ``node``
    (mesh nodes on the nodal basis): NW x NV float64 on the GPU, each row
    normalised to sum 1 over all grid points and negated, then zero at
    points outside the mesh (``inmesh``).  As in the reference, the Gaussian
    width is s = 2*sqrt(2 ln 2) * w and the square spot is the box
    |x - pos| <= w/2.  (On the host this was 1.2 s of a 3.2 s reconstruction
    at 243 spots x 269k nodes.)"""
    node = torch.as_tensor(np.asarray(node, float), dtype=torch.float64, device=DEV)
    pos = torch.as_tensor(np.atleast_2d(np.asarray(pos, float)), dtype=torch.float64, device=DEV)
    if kind == "gauss":
        s = 2.0 * math.sqrt(2.0 * math.log(2.0)) * w
        d2 = torch.zeros((pos.shape[0], node.shape[0]), dtype=torch.float64, device=DEV)
        for a in range(3):                        # per axis: no NW x NV x 3 intermediate
            d2 += (node[None, :, a] - pos[:, a, None]) ** 2
        t = torch.exp(-d2 / (2.0 * s * s))
    elif kind == "square":
        h = w / 2.0
        t = torch.ones((pos.shape[0], node.shape[0]), dtype=torch.bool, device=DEV)
        for a in range(3):
            t &= (node[None, :, a] - pos[:, a, None]).abs() <= h
        t = t.to(torch.float64)
    else:
        raise ValueError("Check spot type is gauss or square")
    t = -(t / t.sum(dim=1, keepdim=True))
    if inmesh is not None:
        t[:, ~torch.as_tensor(np.asarray(inmesh, bool), device=DEV)] = 0.0
    return t


@dataclass
class LMResult:
    mua: np.ndarray
    count: int
    err: list = field(default_factory=list)
    errPerc: list = field(default_factory=list)
    errDel: list = field(default_factory=list)
    lam: list = field(default_factory=list)
    exit: str = ""
    # per update: the branch taken ('accept' / 'step back') and the nodal mua
    # after it
    branch: list = field(default_factory=list)
    mua_hist: list = field(default_factory=list)
    # seconds per iteration: forward+adjoint solve, J-TG and data; the dense
    # algebra (reshape_J, tik_inv)
    t_forward: list = field(default_factory=list)
    t_algebra: list = field(default_factory=list)


def _forward(mesh, gate_fc, mode, tol, method, adj_tol=None):
    """TG data (NM x NG, float64) and the Rytov-normalised TG Jacobian as
    rows x NV float32, rows ordered m + NM*g."""
    f = mesh.solve_point_sources("sd", mode=mode, tol=tol, method=method, adj_tol=adj_tol)
    y = tg_data(mesh, f, gate_fc)                                  # NM x NG
    J, _ = jacobian_tg(mesh, gate_fc, field=f, scale=(1.0 / y).to(torch.float32))
    nv, ng, nm = J.shape
    Jr = J.reshape(nv, ng * nm).t().contiguous()                   # rows x NV
    return y, Jr


def wghtd_lm_recon(mesh, data_m, gates, pos_s=None, lam=(1e2, 1e-6), mode="mixed", tol=1e-8,
                   max_iter=10, spot_fwhm=1.5, verbose=False, method=None, adj_tol=None, forward=None):
    """wghtd_LM_recon(mesh, data_m, gates, pos_s, lam): returns LMResult with
    the reconstructed nodal mua.  data_m: NM x NG measured TG data; gates:
    NG x NT gating functions on the optode's time axis.  mode/method: the
    solver form of the forward+adjoint solves ('plain' is BiCGSTAB only);
    adj_tol: tolerance of the adjoint solves, which feed only the Jacobian
    (default tol; the fused and mixed forms).  forward (optional): a wrapper
    f(mesh, run) -> (data NM x NG, J rows x NV) around run(mesh), the GPU
    forward of this function (J_complex and fc2tg of one batched
    forward+adjoint solve), called once per iteration."""
    opt = mesh.optode
    data_m = np.asarray(data_m, float)
    incl = id_thresh(torch.as_tensor(data_m), 0.8).cpu().numpy().ravel(order="F")   # :25-26
    ym = np.log(data_m.ravel(order="F")[incl])                                     # :28-30
    gate_fc = td2fc(torch.as_tensor(np.asarray(gates, float)), opt.fAxis, opt.tAxis, 1)  # :32
    basis = getattr(mesh, "basis", None)
    spots = None
    if pos_s is not None and len(pos_s):
        if basis is not None:
            t = target_spots(basis.grid, pos_s, spot_fwhm, "gauss", basis.gridinmesh)
            t = t[:, torch.as_tensor(np.asarray(basis.in_idx, np.int64), device=DEV)]
        else:
            t = target_spots(mesh.node, pos_s, spot_fwhm, "gauss")
        spots = t.to(torch.float32)                                                # :35-37
        del t

    def to_grid(v):
        """m2g * v restricted to the unknowns (in-mesh voxels or nodes)."""
        if basis is None:
            return np.asarray(v, float).copy()
        return basis.m2g_in.apply(v).cpu().numpy()

    def to_mesh(xu):
        """g2m * x with x zero outside the mesh (nodal basis: x itself)."""
        if basis is None:
            return xu
        xg = np.zeros(basis.nv)
        xg[basis.in_idx] = xu
        return basis.to_mesh(xg).cpu().numpy()
    rows = torch.as_tensor(np.flatnonzero(incl), device=DEV)
    k = 10 ** 0.25
    lam1, lam2 = lam[0] * k, lam[1]
    res = LMResult(mua=mesh.mua.copy(), count=0)
    y_prev = J_prev = d_mu = None
    while True:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run = lambda m: _forward(m, gate_fc, mode, tol, method, adj_tol)          # noqa: E731
        yc, J = run(mesh) if forward is None else forward(mesh, run)               # :58-61
        torch.cuda.synchronize()
        res.t_forward.append(time.perf_counter() - t0)
        yc = yc.cpu().numpy()
        dc = np.log(yc.ravel(order="F")[incl])                                     # :62-63
        J = J.index_select(0, rows)                                                # :64-65
        t1 = time.perf_counter()
        y = ym - dc                                                                # :67
        if spots is not None:                                                      # :71-77
            w = recon.reshape_J(J, spots, lam2)                                    # NW x rows (FP64)
            J = recon.gemm_nn(w.to(torch.float32), J)
            y = (w @ torch.as_tensor(y, device=DEV)).cpu().numpy()
            wdm = (w @ torch.as_tensor(ym, device=DEV)).cpu().numpy()
        else:
            wdm = ym
        res.err.append(float(np.linalg.norm(y)))                                  # :85
        res.errPerc.append(res.err[-1] / float(np.linalg.norm(wdm)) * 100.0)      # :86
        if verbose:
            print(f"[lm] iteration {res.count}: error {res.errPerc[-1]:.3f}%")
        if res.count > 0:
            res.errDel.append(res.errPerc[-1] - res.errPerc[-2])                  # :88-90
        if (res.count > 1 and np.sign(res.errDel[-1]) == -1
                and (abs(res.errDel[-1]) < 0.12 or res.errPerc[-1] < 1.5)):        # :94-97
            res.exit = "converged"
            break
        res.count += 1
        if res.count == 1 or np.sign(res.errDel[-1]) == -1:                       # :103-107
            G = to_grid(mesh.mua)
            Jg = J * torch.as_tensor(G, dtype=torch.float32, device=DEV)[None, :]  # J .* G'
            lam1 = lam1 / k
            res.branch.append("accept")
        else:                                                                       # :108-116
            # the previous values: y and the J .* G' of the last accepted
            # step; G (of the accepted mua) only rescales this update
            G = to_grid(mesh.mua - d_mu)
            y, Jg = y_prev, J_prev
            lam1 = lam1 * k
            res.branch.append("step back")
        if lam1 < res.err[-1]:                                                     # :119-121
            lam1 = res.err[-1]
        res.lam.append(lam1)
        upd = recon.tik_inv(Jg, lam1, y).cpu().numpy()                             # :128
        res.t_algebra.append(time.perf_counter() - t1)
        d_mu = to_mesh(G * upd)                                                    # :130-134
        if np.any(np.isnan(d_mu)):                                                 # :137-141
            res.exit = "nan"
            break
        new_mu = mesh.mua + d_mu                                                   # :143
        if np.any(new_mu < 0):                                                     # :147-152
            res.exit = "negative mua"
            break
        mesh.update_properties(np.stack([new_mu, mesh.musp], 1))                   # :155
        res.mua_hist.append(mesh.mua.copy())
        if res.count > max_iter - 1:                                               # :159-162
            res.exit = "max iterations"
            break
        y_prev, J_prev = y, Jg                                                     # :165-166
    res.mua = mesh.mua.copy()
    return res


def simulate_tg_data(mesh, gates, mode="fp64", tol=1e-10, peak_counts=None, seed=0, method=None):
    """Time-gated boundary data of ``mesh`` (NM x NG): fc2tg(meas_flu(phi))
    of a forward solve, with Poisson noise when ``peak_counts`` is given (the
    data scaled so that their maximum is peak_counts photons, drawn, and
    scaled back)."""
    from .transforms import td2fc
    opt = mesh.optode
    f = mesh.solve_point_sources("s", mode=mode, tol=tol, method=method)
    gate_fc = td2fc(torch.as_tensor(np.asarray(gates, float)), opt.fAxis, opt.tAxis, 1)
    y = tg_data(mesh, f, gate_fc).cpu().numpy()
    del f
    if peak_counts is None:
        return y
    scale = peak_counts / y.max()
    rng = np.random.default_rng(seed)
    counts = rng.poisson(np.clip(y * scale, 0.0, None)).astype(float)
    return counts / scale
