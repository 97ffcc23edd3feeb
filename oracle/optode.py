"""Oracle: optode placement, source vectors and detection (test infrastructure only).

Restates ``DOGPUP/dOptode.m:77-97,155-199`` (snap2mesh / snap_pos),
``dMesh.m:106-129,174-196`` (source vectors of flu_solve / adj_flu_solve) and
``dMesh.m:144-171`` (meas_flu).  Point location is a brute-force barycentric
search over every tetrahedron (MATLAB's pointLocation tie-breaking is not
reproducible and does not change the resulting nodal weights).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .fem import OMesh


def locate(m: OMesh, pts):
    """For each point: (element id, 4 barycentric weights), brute force."""
    node, elem = m.node, m.elem
    out_id = np.full(len(pts), -1, np.int64)
    out_w = np.full((len(pts), 4), np.nan)
    v0 = node[elem[:, 0]]
    T = np.stack([node[elem[:, 1]] - v0, node[elem[:, 2]] - v0, node[elem[:, 3]] - v0], axis=2)
    Tinv = np.linalg.inv(T)
    for i, p in enumerate(np.asarray(pts, float)):
        lam = np.einsum("eij,ej->ei", Tinv, p[None, :] - v0)
        w = np.concatenate([1.0 - lam.sum(axis=1, keepdims=True), lam], axis=1)
        score = w.min(axis=1)
        e = int(np.argmax(score))
        if score[e] > -1e-9:
            out_id[i] = e
            out_w[i] = w[e]
    return out_id, out_w


def _closest_on_triangle(p, a, b, c):
    """Closest point to p on triangle abc (Voronoi-region method)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    denom = 1.0 / (va + vb + vc)
    return a + ab * (vb * denom) + ac * (vc * denom)


def _clean(w):
    w = w.copy()
    w[w < 1e-10] = 0.0
    return w / w.sum(axis=1, keepdims=True)


def snap_pos(m: OMesh, pos):
    """dOptode.m:155-199: nearest boundary node -> nearest point on its faces ->
    barycentric coordinates; inward normal of the chosen face."""
    pos = np.array(pos, float)
    bidx = np.flatnonzero(m.bnd == 1)
    normals = np.zeros_like(pos)
    for i in range(len(pos)):
        nb = bidx[np.argmin(np.linalg.norm(m.node[bidx] - pos[i], axis=1))]
        fi = np.flatnonzero(np.any(m.face == nb, axis=1))
        best, bp = np.inf, None
        for f in fi:
            tri = m.node[m.face[f]]
            q = _closest_on_triangle(pos[i], tri[0], tri[1], tri[2])
            dd = np.linalg.norm(q - pos[i])
            if dd < best:
                best, bp, bf = dd, q, f
        pos[i] = bp
        tri = m.node[m.face[bf]]
        n = -np.cross(tri[1] - tri[0], tri[2] - tri[0])
        normals[i] = n / np.linalg.norm(n)
    eid, w = locate(m, pos)
    return pos, eid, _clean(w), normals


def snap2mesh(m: OMesh, s_pos, d_pos):
    """dOptode.m:77-97: detectors on the surface; sources moved 1/min(musp)
    along the inward normal."""
    d_pos, d_id, d_w, _ = snap_pos(m, d_pos)
    s_pos, s_id, s_w, s_dir = snap_pos(m, s_pos)
    s_in = s_pos + (1.0 / m.musp.min()) * s_dir
    s_id, s_w = locate(m, s_in)
    return dict(s_pos=s_in, s_id=s_id, s_w=_clean(s_w), d_pos=d_pos, d_id=d_id, d_w=d_w)


def source_matrix(m: OMesh, eid, w, spectrum):
    """Q (NN x NF x NS): Q(elem(eid),:,s) = w, times the source spectrum
    (dMesh.m:124-129; the adjoint uses detector weights, :191-196)."""
    nn = m.node.shape[0]
    ns = len(eid)
    Q = np.zeros((nn, 1, ns))
    for s in range(ns):
        Q[m.elem[eid[s]], 0, s] = w[s]
    return Q * np.asarray(spectrum)[None, :, None]


def detection_matrix(m: OMesh, d_id, d_w):
    """D (ND x NN) from detector barycentric weights (dMesh.m:154-162)."""
    nd = len(d_id)
    rows = np.repeat(np.arange(nd), 4)
    cols = m.elem[d_id].ravel()
    vals = d_w.ravel()
    keep = vals > 0
    return sp.csr_matrix((vals[keep], (rows[keep], cols[keep])), shape=(nd, m.node.shape[0]))


def meas_flu(m: OMesh, d_id, d_w, link, phi):
    """data(m,k) = (D phi_k)(d_m, s_m) (dMesh.m:144-171).  link is 0-based
    [source, detector]."""
    D = detection_matrix(m, d_id, d_w)
    nf = phi.shape[1]
    data = np.zeros((link.shape[0], nf), complex)
    for k in range(nf):
        temp = D @ phi[:, k, :]
        data[:, k] = temp[link[:, 1], link[:, 0]]
    return data
