"""Time-gated adjoint absorption Jacobian (J_complex followed by fc2tg).

``jacobian_tg`` solves forward and adjoint fields in one batched launch
(columns 0..NS-1 sources, NS..NS+ND-1 detectors; dMesh.m:174-202 builds the
adjoint sources from the detector weights times s_fpsf) and contracts them
with the gate spectra in one kernel:

    J_TG(m, g, v) = sum_k Re(C_gk phi(v,k,s_m) phiA(v,k,d_m)) * scale(m, g)
    C_gk = -2 Nt w_k conj(Gate_gk) dxyz^3 / s_fpsf(k),  w_0 = 1/2,

which equals ``fc2tg(J_complex(...), gates, f, Nt, 2)`` (dMesh.m:206-247,
fc2tg.m:25-32) on the nodal basis (grid = lattice nodes, m2g = I).
``AbsorptionTG`` exposes the map mua -> time-gated data with this Jacobian as
its derivative, as a torch.autograd.Function.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import _lib
from .transforms import td2fc

DEV = "cuda"


def boxcar_gates(t, n_gates=8, width=500e-12, spacing=200e-12, start=0.0):
    g = np.zeros((n_gates, t.size))
    for i in range(n_gates):
        t0 = start + i * spacing
        g[i] = ((t >= t0) & (t < t0 + width)).astype(float)
    return g


def boxcar_gates_fc(optode, t, n_gates=8, width=500e-12, spacing=200e-12):
    """Gate spectra td2fc(gates) (NG x NF complex, on the device)."""
    return td2fc(torch.as_tensor(boxcar_gates(np.asarray(t), n_gates, width, spacing)),
                 optode.fAxis, t, 1)


def tg_coefficients(optode, gate_fc, dxyz):
    nf = optode.Nf
    w = np.where(np.arange(nf) == 0, 0.5, 1.0)
    g = gate_fc.cpu().numpy() if torch.is_tensor(gate_fc) else np.asarray(gate_fc)
    coef = -2.0 * optode.tAxis.size * w[None, :] * np.conj(g) * dxyz ** 3 / optode.s_fpsf[None, :]
    return torch.as_tensor(np.ascontiguousarray(coef), dtype=torch.complex128, device=DEV)


def jacobian_tg(mesh, gate_fc, mode="mixed", tol=1e-10, dxyz=None, scale=None, field=None):
    """J_TG as float32 [NV][NG][NM] (= the reference's NM x NG x NV,
    column-major).  ``field``: a forward+adjoint FluenceField from
    ``mesh.solve_point_sources('sd')`` to reuse; ``scale``: optional NM x NG
    factor (e.g. 1/data for the Rytov normalisation)."""
    opt = mesh.optode
    ns, nd = len(opt.s_bary[0]), len(opt.d_bary[0])
    if dxyz is None:
        dxyz = getattr(mesh, "dxyz", None)
        if dxyz is None:
            raise ValueError("voxel side dxyz is needed (set mesh.dxyz or pass dxyz)")
    t0 = time.perf_counter()
    if field is None:
        field = mesh.solve_point_sources("sd", mode=mode, tol=tol)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    coef = tg_coefficients(opt, gate_fc, dxyz)
    ng, nm, nn = coef.shape[0], opt.link.shape[0], mesh.nn
    link = torch.as_tensor(np.stack([opt.link[:, 0], opt.link[:, 1] + ns], 1).astype(np.int32), device=DEV)
    out = torch.empty((nn, ng, nm), dtype=torch.float32, device=DEV)
    sc = None
    if scale is not None:
        sc = torch.as_tensor(scale, dtype=torch.float32, device=DEV).t().contiguous()   # NG x NM
    x = torch.view_as_real(field.data)
    _lib.call("dogpup_jacobian_tg", _lib.ptr(x), _lib.ptr(torch.view_as_real(coef)), _lib.ptr(link),
              _lib.ptr(sc), _lib.ptr(out), nn, opt.Nf, ng, field.ncols, field.nsp, nm, nn,
              _lib.stream_ptr())
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return out, {"solve_s": t1 - t0, "kernel_s": t2 - t1, "field": field, "coef": coef, "link": link}
