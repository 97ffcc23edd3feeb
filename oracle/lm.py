"""Oracle: weighted Levenberg-Marquardt absorption reconstruction (test
infrastructure only).

``wghtd_lm_recon`` restates ``DOGPUP/reconstruction/wghtd_LM_recon.m:23-168``
on the nodal basis (grid = mesh nodes: m2g = g2m = I, gridinmesh all true),
with ``target_spots`` (``dMesh.m:250-282``), ``id_thresh`` (``id_thresh.m``),
``reshape_J`` and ``tik_inv`` (``reconstruction/*.m``) in FP64.  The forward
model is a callable ``forward(mua) -> (data_tg NM x NG, J_tg NM x NG x NV)``
so that the loop can run on the FP64 FEM restatement (``fem_forward``:
direct sparse solves, J_complex, fc2tg and the Rytov division,
wghtd_LM_recon.m:58-61) or on a closed-form model.

Every iteration is recorded (error, the branch taken, lambda, mua), so that
a check can compare whole trajectories, step-back branches included.
"""
from __future__ import annotations

import math

import numpy as np

from . import fem, jacobian as ojac, optode as oopt, recon as orec, solver as osol, transforms as otr


def target_spots(grid, pos, w, kind="gauss", gridinmesh=None):
    """dMesh.m:250-282: NS x NV spots, each normalised to sum 1 over the grid
    and negated, zero outside the mesh.  As the reference does, the Gaussian
    width is s = 2 sqrt(2 ln 2) w (so the spot's half-maximum width is
    8 ln 2 w), the square spot is the box |x - pos| <= w/2."""
    grid = np.asarray(grid, float)
    P = np.atleast_2d(np.asarray(pos, float)).T                 # 3 x NS (:262)
    if kind == "gauss":
        s = (2.0 * math.sqrt(2.0 * math.log(2.0))) * w          # :264
        t = np.exp(-(grid[:, 0:1] - P[0]) ** 2 / (s ** 2 * 2)
                   - (grid[:, 1:2] - P[1]) ** 2 / (s ** 2 * 2)
                   - (grid[:, 2:3] - P[2]) ** 2 / (s ** 2 * 2))  # NV x NS (:265-267)
    elif kind == "square":
        h = w / 2.0
        t = np.ones((grid.shape[0], P.shape[1]), bool)
        for a in range(3):
            t &= (grid[:, a:a + 1] <= P[a] + h) & (grid[:, a:a + 1] >= P[a] - h)
        t = t.astype(float)
    else:
        raise ValueError("Check spot type is gauss or square")
    t = t / t.sum(axis=0)                                        # :278
    t = -t.T                                                     # :279
    if gridinmesh is not None:
        t[:, ~np.asarray(gridinmesh, bool)] = 0.0                # :280
    return t


def fem_forward(m: fem.OMesh, snap, link, s_fpsf, gate_fc, f, df, nt, dxyz):
    """forward(mua) of the FP64 FEM restatement: for nodal absorption mua,
    update_properties (dMesh.m:80-88, kappa follows mua), the forward and
    adjoint fields by direct sparse solves (flu_solve / adj_flu_solve,
    dMesh.m:106-202 to convergence), the time-gated data fc2tg(meas_flu) and
    J = fc2tg(J_complex) ./ data (wghtd_LM_recon.m:58-61)."""
    def forward(mua):
        fem.update_properties(m, mua, m.musp)
        A = fem.full_system(m, df)
        phi = osol.direct(A, oopt.source_matrix(m, snap["s_id"], snap["s_w"], s_fpsf))
        phiA = osol.direct(A, oopt.source_matrix(m, snap["d_id"], snap["d_w"], s_fpsf))
        data = oopt.meas_flu(m, snap["d_id"], snap["d_w"], link, phi)             # NM x NF
        data_tg = otr.fc2tg(data, gate_fc, f, nt, 1)                              # NM x NG
        J = ojac.j_complex(phi, phiA, link, s_fpsf, dxyz)                          # NM x NF x NV
        J_tg = otr.fc2tg(J, gate_fc, f, nt, 1) / data_tg[:, :, None]
        return data_tg, J_tg
    return forward


def wghtd_lm_recon(forward, mua0, data_m, spots=None, lam=(1e2, 1e-6), max_iter=10):
    """wghtd_LM_recon.m:23-168 with the nodal basis.  forward(mua) gives the
    time-gated data and the Rytov-normalised Jacobian; data_m: NM x NG;
    spots: NW x NV target spots or None.  Returns a dict with the final mua,
    count, err, errPerc, errDel, the exit and, per update, the branch taken
    ('accept' or 'step back'), lambda and the mua after the update."""
    mua = np.asarray(mua0, float).copy()
    data_m = np.asarray(data_m, float)
    id_incl = otr.id_thresh(data_m, 0.8).ravel(order="F")                    # :26-27
    # non-positive gated values (band-limited ringing before the arrival)
    # have no real log; the selection drops them, as it drops the complex
    # logs MATLAB forms for them
    with np.errstate(invalid="ignore", divide="ignore"):
        ym = np.log(data_m.ravel(order="F"))[id_incl]                        # :29-30
    k = 10.0 ** 0.25                                                         # :40
    lam1, lam2 = lam[0] * k, lam[1]                                          # :41
    count = 0
    err, errPerc, errDel = [], [], []
    out = {"branch": [], "lam": [], "mua_hist": [], "exit": ""}
    y_prev = J_prev = d_mu = None
    while True:
        data_c, J = forward(mua)                                             # :58-61
        with np.errstate(invalid="ignore", divide="ignore"):
            dc = np.log(np.asarray(data_c, float).ravel(order="F"))[id_incl]  # :62-63
        nv = J.shape[2]
        J = np.asarray(J, float).reshape(-1, nv, order="F")[id_incl]         # :64-65
        y = ym - dc                                                          # :67
        if spots is not None:                                                # :71-77
            w = orec.reshape_J(J, spots, lam2)
            J = w @ J
            y = w @ y
            wdm = w @ ym
        else:
            wdm = ym
        err.append(float(np.linalg.norm(y)))                                 # :85
        errPerc.append(err[-1] / float(np.linalg.norm(wdm)) * 100.0)         # :86
        if count > 0:
            errDel.append(errPerc[-1] - errPerc[-2])                         # :88-90
        if count > 1 and np.sign(errDel[-1]) == -1 and (abs(errDel[-1]) < 0.12 or errPerc[-1] < 1.5):
            out["exit"] = "converged"                                        # :94-97
            break
        count += 1                                                           # :99
        if count == 1 or np.sign(errDel[-1]) == -1:                          # :103-107
            G = mua.copy()
            J = J * G[None, :]
            lam1 = lam1 / k
            out["branch"].append("accept")
        else:                                                                # :108-116
            G = mua - d_mu
            y = y_prev
            J = J_prev
            lam1 = lam1 * k
            out["branch"].append("step back")
        if lam1 < err[-1]:                                                   # :119-121
            lam1 = err[-1]
        out["lam"].append(lam1)
        update = orec.tik_inv(J, lam1, y)                                    # :128
        d_mu = G * update                                                    # :130-134
        if np.any(np.isnan(d_mu)):                                           # :137-141
            out["exit"] = "nan"
            break
        new_mu = mua + d_mu                                                  # :143
        if np.any(new_mu < 0):                                               # :147-152
            out["exit"] = "negative mua"
            break
        mua = new_mu                                                         # :155
        out["mua_hist"].append(mua.copy())
        if count > max_iter - 1:                                             # :159-162
            out["exit"] = "max iterations"
            break
        y_prev = y                                                           # :165-166
        J_prev = J
    out.update(mua=mua, count=count, err=err, errPerc=errPerc, errDel=errDel)
    return out
