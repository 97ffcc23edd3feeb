"""Time-gated adjoint absorption Jacobian (J_complex followed by fc2tg).

``jacobian_tg`` solves forward and adjoint fields in one batched launch
(columns 0..NS-1 sources, NS..NS+ND-1 detectors; dMesh.m:174-202 builds the
adjoint sources from the detector weights times s_fpsf) and contracts them
with the gate spectra in one kernel:

    J_TG(m, g, v) = sum_k Re(C_gk phi(v,k,s_m) phiA(v,k,d_m)) * scale(m, g)
    C_gk = -2 Nt w_k conj(Gate_gk) dxyz^3 / s_fpsf(k),  w_0 = 1/2,

which equals ``fc2tg(J_complex(...), gates, f, Nt, 2)`` (dMesh.m:206-247,
fc2tg.m:25-32).  On the nodal basis (no voxel grid attached: grid = mesh
nodes, m2g = I) v runs over the nodes; with a grid from ``mesh.mesh2grid`` the
fields are first interpolated to the in-mesh voxels (m2g, dMesh.m:239-240)
and v runs over those (``mesh.basis.in_idx``; the reference's columns outside
the mesh are zero).
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


def jacobian_tg(mesh, gate_fc, mode="mixed", tol=1e-10, dxyz=None, scale=None, field=None,
                method=None, adj_tol=None):
    """J_TG as float32 [NV][NG][NM] (= the reference's NM x NG x NV,
    column-major); NV = nodes, or in-mesh voxels when ``mesh.basis`` is set.
    ``field``: a forward+adjoint FluenceField from
    ``mesh.solve_point_sources('sd')`` to reuse; ``scale``: optional NM x NG
    factor (e.g. 1/data for the Rytov normalisation); ``adj_tol``: tolerance
    of the adjoint (detector) columns when it differs from ``tol``."""
    opt = mesh.optode
    ns, nd = len(opt.s_bary[0]), len(opt.d_bary[0])
    basis = getattr(mesh, "basis", None)
    if dxyz is None:
        dxyz = basis.dxyz if basis is not None else getattr(mesh, "dxyz", None)
        if dxyz is None:
            raise ValueError("voxel side dxyz is needed (set mesh.dxyz or pass dxyz)")
    t0 = time.perf_counter()
    if field is None:
        field = mesh.solve_point_sources("sd", mode=mode, tol=tol, method=method, adj_tol=adj_tol)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    vfield = basis.field_to_voxels(field) if basis is not None else field
    coef = tg_coefficients(opt, gate_fc, dxyz)
    ng, nm, nn = coef.shape[0], opt.link.shape[0], vfield.nn
    link = torch.as_tensor(np.stack([opt.link[:, 0], opt.link[:, 1] + ns], 1).astype(np.int32), device=DEV)
    out = torch.empty((nn, ng, nm), dtype=torch.float32, device=DEV)
    sc = None
    if scale is not None:
        sc = torch.as_tensor(scale, dtype=torch.float32, device=DEV).t().contiguous()   # NG x NM
    x = torch.view_as_real(vfield.data)
    _lib.call("dogpup_jacobian_tg", _lib.ptr(x), _lib.ptr(torch.view_as_real(coef)), _lib.ptr(link),
              _lib.ptr(sc), _lib.ptr(out), nn, opt.Nf, ng, vfield.ncols, vfield.nsp, nm, nn,
              _lib.stream_ptr())
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return out, {"solve_s": t1 - t0, "kernel_s": t2 - t1, "field": field, "coef": coef, "link": link}


def _links_by_source(link, ns):
    order = np.argsort(link[:, 0], kind="stable").astype(np.int32)
    ptr = np.zeros(ns + 1, np.int64)
    np.add.at(ptr, link[:, 0] + 1, 1)
    return np.cumsum(ptr).astype(np.int32), order


def tg_data(mesh, field, gate_fc):
    """Time-gated boundary data y = fc2tg(meas_flu(phi)) (NM x NG, FP64)."""
    from .transforms import fc2tg
    data = mesh.meas_flu(field)                           # NM x NF complex
    return fc2tg(data, gate_fc, mesh.optode.fAxis, mesh.optode.tAxis.size, 1)


def vjp(mesh, field, gate_fc, ybar, exact=False, dxyz=None):
    """g = J_TG^T ybar for nodal mua (kappa fixed), J not formed.  exact=False
    is the reference's nodal-product Jacobian on the nodal basis, scaled by
    dxyz^3 as jacobian_tg scales it (dxyz: the argument, else mesh.dxyz, else
    ValueError); exact=True the exact FEM derivative, which has no voxel
    factor.  The nodal-product form is rejected on a mesh with a voxel basis
    attached: its J_TG has voxel columns, of which a nodal gradient is not the
    transpose."""
    if not exact:
        if getattr(mesh, "basis", None) is not None:
            raise ValueError("vjp: the nodal-product J_TG of a mesh with a voxel basis has voxel columns; "
                             "use jacobian_tg, or clear_grid() for the nodal basis")
        if dxyz is None:
            dxyz = getattr(mesh, "dxyz", None)
        if dxyz is None:
            raise ValueError("vjp: voxel side dxyz is needed (set mesh.dxyz or pass dxyz)")
    else:
        dxyz = 1.0                                    # unused by the exact derivative
    opt = mesh.optode
    ns = len(opt.s_bary[0])
    nf, nt = opt.Nf, opt.tAxis.size
    g = gate_fc.cpu().numpy() if torch.is_tensor(gate_fc) else np.asarray(gate_fc)
    w = np.where(np.arange(nf) == 0, 0.5, 1.0)
    cp = 2.0 * nt * w[None, :] * np.conj(g) / opt.s_fpsf[None, :]          # NG x NF
    yb = ybar.detach().cpu().numpy() if torch.is_tensor(ybar) else np.asarray(ybar)
    D = np.ascontiguousarray((yb @ cp).T)                                 # NF x NM
    Dd = torch.as_tensor(D, dtype=torch.complex128, device=DEV)
    link = np.stack([opt.link[:, 0], opt.link[:, 1] + ns], 1).astype(np.int32)
    lptr, lidx = _links_by_source(opt.link, ns)
    link_d, lptr_d, lidx_d = (torch.as_tensor(a, device=DEV) for a in (link, lptr, lidx))
    E = torch.empty((mesh.nn, nf, ns), dtype=torch.complex128, device=DEV)
    out = torch.empty(mesh.nn, dtype=torch.float64, device=DEV)
    x = torch.view_as_real(field.data)
    _lib.call("dogpup_jacobian_vjp", _lib.ptr(x), _lib.ptr(torch.view_as_real(Dd)), _lib.ptr(link_d),
              _lib.ptr(lptr_d), _lib.ptr(lidx_d), _lib.ptr(mesh._dev["node"]), _lib.ptr(mesh._dev["elem"]),
              _lib.ptr(torch.view_as_real(E)), _lib.ptr(out), mesh.nn, mesh.elem.shape[0], nf, field.ncols,
              field.nsp, ns, opt.link.shape[0], int(exact), float(dxyz) ** 3, _lib.stream_ptr())
    return out


class AbsorptionTG(torch.autograd.Function):
    """mua (NN, float64) -> time-gated boundary data (NM x NG) with kappa held
    at the mesh's current values; backward = J_TG^T ybar through
    dogpup_jacobian_vjp (nodal-product Jacobian of dMesh.m:206-247 with the
    voxel side dxyz, or the exact derivative with exact=True).  The mesh's
    properties (mua, musp, kappa and the assembled system) are restored after
    the forward solve, so the call leaves the caller's mesh as it found it."""

    @staticmethod
    def forward(ctx, mua, mesh, gate_fc, exact=False, mode="fp64", tol=1e-12, dxyz=None):
        props0 = np.stack([mesh.mua.copy(), mesh.musp.copy()], 1)
        kappa0 = mesh.kappa.copy()
        mesh.update_properties(np.stack([mua.detach().cpu().numpy(), props0[:, 1]], 1), kappa=kappa0)
        try:
            field = mesh.solve_point_sources("sd", mode=mode, tol=tol)
            y = tg_data(mesh, field, gate_fc)
        finally:
            mesh.update_properties(props0, kappa=kappa0)
        ctx.mesh, ctx.field, ctx.gate_fc, ctx.exact, ctx.dxyz = mesh, field, gate_fc, exact, dxyz
        ctx.in_device, ctx.in_dtype = mua.device, mua.dtype
        return y

    @staticmethod
    def backward(ctx, ybar):
        g = vjp(ctx.mesh, ctx.field, ctx.gate_fc, ybar, exact=ctx.exact, dxyz=ctx.dxyz)
        return g.to(ctx.in_device, ctx.in_dtype), None, None, None, None, None, None


def absorption_tg(mua, mesh, gate_fc, exact=False, mode="fp64", tol=1e-12, dxyz=None):
    return AbsorptionTG.apply(mua, mesh, gate_fc, exact, mode, tol, dxyz)
