"""Oracle: voxel-grid basis and mesh-to-mesh interpolation (test
infrastructure only).

Restates ``DOGPUP/dMesh.m:286-380`` (mesh2grid) and ``dMesh.m:383-411``
(mesh2mesh) with brute-force point location: every query point is tested
against every tetrahedron (``optode.locate``), and the grid's triangulation
is written out tet by tet.

The reference triangulates the voxel grid with MATLAB's ``delaunay``; a box
grid is degenerate for Delaunay (every cell's 8 corners are cospherical), so
any split of the cells into tetrahedra with those corners is a Delaunay
triangulation and MATLAB's choice among them is not knowable here (parity
unpinned for nodes strictly inside a cell).  The oracle uses the Kuhn split:
six tetrahedra per cell, one per permutation p of the axes, with vertices
0, e_p0, e_p0+e_p1, (1,1,1), enumerated in lexicographic order of p.  Tests
pin the split's Delaunay property (no grid point inside any circumsphere) and
the invariants that hold for every triangulation (linear reproduction, rows
summing to 1, identity on grid points).

``mesh2mesh`` fills points outside the old mesh from the nearest located
point; the reference intends this (dMesh.m:405-409) but indexes node
coordinates instead of node ids there, which only works when every point is
located.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp


def _locate(node, elem, pts, tol=1e-9):
    """Containing tetrahedron (largest smallest weight, first on ties) and
    barycentric weights, brute force over every tetrahedron."""
    node = np.asarray(node, float)
    elem = np.asarray(elem, np.int64)
    pts = np.atleast_2d(np.asarray(pts, float))
    v0 = node[elem[:, 0]]
    T = np.stack([node[elem[:, 1]] - v0, node[elem[:, 2]] - v0, node[elem[:, 3]] - v0], axis=2)
    Tinv = np.linalg.inv(T)
    eid = np.full(len(pts), -1, np.int64)
    w = np.full((len(pts), 4), np.nan)
    for i, p in enumerate(pts):
        lam = np.einsum("eij,ej->ei", Tinv, p[None, :] - v0)
        ww = np.concatenate([1.0 - lam.sum(axis=1, keepdims=True), lam], axis=1)
        score = ww.min(axis=1)
        e = int(np.argmax(score))
        if score[e] >= -tol:
            eid[i] = e
            w[i] = ww[e]
    return eid, w


def ndgrid_points(x, y, z):
    """[X(:), Y(:), Z(:)] of ndgrid(x, y, z): x varies fastest."""
    X, Y, Z = np.meshgrid(np.asarray(x, float), np.asarray(y, float), np.asarray(z, float), indexing="ij")
    return np.stack([X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")], axis=1)


def kuhn_tets(nx, ny, nz):
    """Tetrahedra (voxel ids, voxel v = ix + nx*(iy + ny*iz)) of the Kuhn
    split of an nx x ny x nz grid: cells in voxel order, six per cell."""
    perms = list(itertools.permutations(range(3)))          # lexicographic
    ix, iy, iz = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    cells = np.stack([ix.ravel(order="F"), iy.ravel(order="F"), iz.ravel(order="F")], axis=1)
    out = []
    for c in cells:
        for p in perms:
            off = np.zeros(3, np.int64)
            verts = [c.copy()]
            for a in p:
                off[a] = 1
                verts.append(c + off)
            out.append([v[0] + nx * (v[1] + ny * v[2]) for v in verts])
    return np.asarray(out, np.int64)


def grid_to_mesh(node, x, y, z, inmesh, tol=1e-9):
    """g2m (NN x NV, CSR) from the grid's Kuhn split (dMesh.m:331-373):
    weights on out-of-mesh voxels dropped, nodes with no in-mesh vertex take
    the nearest in-mesh voxel with weight 1, rows divided by their sum (a zero
    sum left as is).  Also returns the fallback node ids."""
    grid = ndgrid_points(x, y, z)
    nx, ny, nz = len(x), len(y), len(z)
    elemG = kuhn_tets(nx, ny, nz)
    gid, gb = _locate(grid, elemG, node, tol)
    if (gid < 0).any():
        raise ValueError("nodes outside the voxel grid")
    nn = len(node)
    j = elemG[gid]                                           # NN x 4
    ein = np.asarray(inmesh, bool)[j]
    rows = np.repeat(np.arange(nn), 4).reshape(nn, 4)
    i_ = rows[ein]
    j_ = j[ein]
    v_ = gb[ein]
    fb = np.flatnonzero(ein.sum(axis=1) < 1)
    if fb.size:
        vin = np.flatnonzero(inmesh)
        d2 = ((node[fb][:, None, :] - grid[vin][None, :, :]) ** 2).sum(axis=2)
        jS = vin[np.argmin(d2, axis=1)]
        i_ = np.concatenate([i_, fb])
        j_ = np.concatenate([j_, jS])
        v_ = np.concatenate([v_, np.ones(fb.size)])
    g2m = sp.csr_matrix((v_, (i_, j_)), shape=(nn, grid.shape[0]))
    s = np.asarray(g2m.sum(axis=1)).ravel()
    s[s == 0] = 1.0
    g2m = sp.diags(1.0 / s) @ g2m
    return sp.csr_matrix(g2m), fb


def mesh2grid(node, elem, x, y, z, tol=1e-9):
    """dict(grid, gridSize, dxyz, m2g (NV x NN), g2m (NN x NV), gridinmesh)."""
    node = np.asarray(node, float)
    elem = np.asarray(elem, np.int64)
    grid = ndgrid_points(x, y, z)
    eid, bary = _locate(node, elem, grid, tol)
    inmesh = eid >= 0
    i = np.repeat(np.flatnonzero(inmesh), 4)
    j = elem[eid[inmesh]].ravel()
    v = bary[inmesh].ravel()
    m2g = sp.csr_matrix((v, (i, j)), shape=(grid.shape[0], node.shape[0]))
    g2m, fb = grid_to_mesh(node, x, y, z, inmesh, tol)
    return dict(grid=grid, gridSize=(len(x), len(y), len(z)), dxyz=abs(x[1] - x[0]), m2g=m2g, g2m=g2m,
                gridinmesh=inmesh, fallback=fb)


def mesh2mesh(node, val, old_node, old_elem, tol=1e-9):
    """Values ``val`` (old NN x C) interpolated to ``node``; points outside the
    old mesh take the value of the nearest located point."""
    val = np.asarray(val, float).reshape(len(old_node), -1)
    eid, w = _locate(old_node, old_elem, node, tol)
    out = np.zeros((len(node), val.shape[1]))
    ok = eid >= 0
    out[ok] = np.einsum("pq,pqc->pc", w[ok], val[np.asarray(old_elem)[eid[ok]]])
    bad = np.flatnonzero(~ok)
    if bad.size:
        good = np.flatnonzero(ok)
        d2 = ((np.asarray(node)[bad][:, None, :] - np.asarray(node)[good][None, :, :]) ** 2).sum(axis=2)
        out[bad] = out[good[np.argmin(d2, axis=1)]]
    return out
