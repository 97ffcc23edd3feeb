"""Oracle: adjoint absorption Jacobian (test infrastructure only).

``j_complex`` restates ``DOGPUP/dMesh.m:206-247`` with the nodal basis
(m2g = I, grid = lattice nodes, voxel side dxyz):
    J(m,k,v) = -phi(v,k,s_m) * phiA(v,k,d_m) * dxyz^3 / s_fpsf(k).
``j_exact`` is the exact FEM derivative of meas_flu with respect to nodal mua
at fixed kappa (the quantity central finite differences measure):
    dJ(m,k,v) = -sum_ij phiA(i,k,d_m) C_v(i,j) phi(j,k,s_m) / s_fpsf(k),
    C_v(i,j) = int N_v N_i N_j.
``j_tg`` composes j_complex with fc2tg along the frequency axis
(wghtd_LM_recon.m:60-61 without the Rytov division).
"""
from __future__ import annotations

import numpy as np

from .fem import OMesh
from . import transforms as otr


def j_complex(phi, phiA, link, s_fpsf, dxyz):
    """phi NN x NF x NS, phiA NN x NF x ND -> J NM x NF x NV (NV = NN)."""
    J = -phi[:, :, link[:, 0]] * phiA[:, :, link[:, 1]] * dxyz ** 3     # NV x NF x NM
    J = np.transpose(J, (2, 1, 0))
    return J / np.asarray(s_fpsf)[None, :, None]


def j_tg(phi, phiA, link, s_fpsf, dxyz, gate_fc, f, nt):
    """fc2tg(J_complex) -> NM x NG x NV."""
    return otr.fc2tg(j_complex(phi, phiA, link, s_fpsf, dxyz), gate_fc, f, nt, 1)


def absorption_mass_column(m: OMesh, v: int):
    """Sparse entries of C_v(i,j) = int N_v N_i N_j over the tets around node v
    (closed-form P1 integrals: 1/20, 1/60, 1/120 of the volume)."""
    tets = np.flatnonzero(np.any(m.elem == v, axis=1))
    ent = {}
    for e in tets:
        nodes = m.elem[e]
        for a in range(4):
            for b in range(4):
                i, j = nodes[a], nodes[b]
                idx = [v, i, j]
                cnt = sorted([idx.count(u) for u in set(idx)], reverse=True)
                c = {(3,): 1 / 20, (2, 1): 1 / 60, (1, 1, 1): 1 / 120}[tuple(cnt)]
                ent[(i, j)] = ent.get((i, j), 0.0) + m.vol[e] * c
    return ent


def j_exact(m: OMesh, phi, phiA, link, s_fpsf, nodes):
    """Exact d(meas_flu)/d(mua_v) at fixed kappa for v in ``nodes``:
    NM x NF x len(nodes)."""
    nm, nf = link.shape[0], phi.shape[1]
    out = np.zeros((nm, nf, len(nodes)), complex)
    for q, v in enumerate(nodes):
        ent = absorption_mass_column(m, v)
        ii = np.array([k[0] for k in ent])
        jj = np.array([k[1] for k in ent])
        cc = np.array(list(ent.values()))
        a = phiA[ii][:, :, link[:, 1]]     # E x NF x NM
        b = phi[jj][:, :, link[:, 0]]
        out[:, :, q] = -np.einsum("e,efm->mf", cc, a * b)
    return out / np.asarray(s_fpsf)[None, :, None]
