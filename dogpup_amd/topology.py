"""Host-side mesh topology: boundary faces, CSR patterns and scatter maps.

This is the integer bookkeeping that the reference does on the host in
MATLAB (``DOGPUP/dMesh.m:945-1030`` prepare_mesh, ``:1129-1140`` face
selection, ``:1351-1414`` FSAIP_gen pattern and sparse_csr).  It runs once
per mesh topology; every floating-point integral runs on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

C_VAC = 2.99792458e11  # mm/s (dMesh.m:42)


def robin_factor(nr: float) -> float:
    """R = 1/(2A) for a medium/air boundary (dMesh.m:1008-1011)."""
    r0 = (nr - 1.0) ** 2 / (nr + 1.0) ** 2
    cos_c = np.cos(np.arcsin(1.0 / nr))
    a = (2.0 / (1.0 - r0) - 1.0 + abs(cos_c) ** 3) / (1.0 - abs(cos_c) ** 2)
    return 1.0 / (2.0 * a)


def _face_keys(f: np.ndarray, nn: int) -> np.ndarray:
    s = np.sort(f, axis=1).astype(np.int64)
    return (s[:, 0] * nn + s[:, 1]) * nn + s[:, 2]


def true_faces(node: np.ndarray, elem: np.ndarray):
    """Faces that belong to one tetrahedron only, oriented outward, and their
    areas (dMesh.m:961-1002)."""
    nn = node.shape[0]
    faces = np.concatenate([elem[:, [0, 1, 2]], elem[:, [0, 1, 3]], elem[:, [0, 2, 3]], elem[:, [1, 2, 3]]])
    keys = _face_keys(faces, nn)
    order = np.argsort(keys, kind="stable")
    ks = keys[order]
    first = np.ones(ks.size, bool)
    first[1:] = ks[1:] != ks[:-1]
    grp = np.cumsum(first) - 1
    counts = np.bincount(grp)
    once = counts[grp] == 1
    face = faces[order[once & first]].astype(np.int64)
    com = node[face].mean(axis=1)
    centre = 0.5 * (node.max(axis=0) + node.min(axis=0))
    vdir = centre[None, :] - com
    vdir /= np.linalg.norm(vdir, axis=1, keepdims=True)
    n = np.cross(node[face[:, 1]] - node[face[:, 0]], node[face[:, 2]] - node[face[:, 0]])
    flip = np.einsum("ij,ij->i", n, vdir) > 0
    face[flip, 1], face[flip, 2] = face[flip, 2].copy(), face[flip, 1].copy()
    n = np.cross(node[face[:, 1]] - node[face[:, 0]], node[face[:, 2]] - node[face[:, 0]])
    area = 0.5 * np.linalg.norm(n, axis=1)
    return face, area


def robin_faces(elem: np.ndarray, bnd: np.ndarray, reference_selection: bool = True, face=None):
    """Faces carrying the Robin integral.  With ``reference_selection`` this is
    the reference's list (dMesh.m:1129-1140): all 4 faces of tetrahedra whose
    nodes are all on the boundary, plus the boundary-node face of tetrahedra
    with exactly 3 boundary nodes; otherwise the true boundary faces."""
    if not reference_selection:
        return np.asarray(face, np.int64)
    srt = np.sort(elem, axis=1).astype(np.int64)
    on = bnd[srt] > 0
    cnt = on.sum(axis=1)
    t4 = srt[cnt == 4]
    f4 = np.concatenate([t4[:, [0, 1, 2]], t4[:, [0, 1, 3]], t4[:, [0, 2, 3]], t4[:, [1, 2, 3]]])
    m3 = cnt == 3
    f3 = srt[m3][on[m3]].reshape(-1, 3)
    return np.concatenate([f4, f3])


@dataclass
class Pattern:
    nn: int
    rptr: np.ndarray        # int32 NN+1, full symmetric A
    cptr: np.ndarray        # int32 nnz
    slot_ptr: np.ndarray    # int32 nnz+1
    slot_src: np.ndarray    # int32 contributions (element / face codes)
    diag: np.ndarray        # int64 NN, slot of (i, i)


def system_pattern(nn: int, elem: np.ndarray, bface: np.ndarray) -> Pattern:
    """CSR pattern of the full symmetric A and, per slot, the element codes
    e*16 + a*4 + b and face codes NE*16 + f*9 + a*3 + b that sum into it."""
    ne = elem.shape[0]
    e64 = elem.astype(np.int64)
    I = np.repeat(e64[:, :, None], 4, axis=2).reshape(-1)     # (e, a, b) -> elem[e, a]
    J = np.repeat(e64[:, None, :], 4, axis=1).reshape(-1)     # (e, a, b) -> elem[e, b]
    keys = I * nn + J
    uk, inv = np.unique(keys, return_inverse=True)
    rows = uk // nn
    cols = (uk % nn).astype(np.int32)
    rptr = np.zeros(nn + 1, np.int64)
    np.add.at(rptr, rows + 1, 1)
    rptr = np.cumsum(rptr)
    slot_e = inv.astype(np.int64)
    codes_e = np.arange(16 * ne, dtype=np.int64)
    if bface.shape[0]:
        bf = bface.astype(np.int64)
        Ib = np.repeat(bf[:, :, None], 3, axis=2).reshape(-1)
        Jb = np.repeat(bf[:, None, :], 3, axis=1).reshape(-1)
        slot_f = np.searchsorted(uk, Ib * nn + Jb)
        codes_f = 16 * ne + np.arange(9 * bf.shape[0], dtype=np.int64)
        slots = np.concatenate([slot_e, slot_f])
        codes = np.concatenate([codes_e, codes_f])
    else:
        slots, codes = slot_e, codes_e
    order = np.argsort(slots, kind="stable")
    slot_src = codes[order]
    slot_ptr = np.zeros(uk.size + 1, np.int64)
    np.add.at(slot_ptr, slots + 1, 1)
    slot_ptr = np.cumsum(slot_ptr)
    diag = np.searchsorted(uk, np.arange(nn, dtype=np.int64) * (nn + 1))
    if slot_src.max() >= 2 ** 31 - 1:
        raise ValueError("mesh too large for int32 contribution codes")
    return Pattern(nn, rptr.astype(np.int32), cols, slot_ptr.astype(np.int32),
                   slot_src.astype(np.int32), diag)


def fsai_pattern(rptr: np.ndarray, cptr: np.ndarray, absval: np.ndarray | None = None, N: int = 30):
    """Pattern of G from tril(A) (FSAIP_gen, dMesh.m:1361-1375).

    In the reference's order (rows descending, |a| descending within a row)
    the kept positions are each row's first entry and the N-1 positions
    before it.  Every row with at most N lower entries therefore keeps all of
    them; longer rows keep their largest entry plus their N-1 smallest (and
    whatever the window of the next row spills over).  ``absval`` (|A| at
    unit frequency, on the full pattern) is needed only when a row exceeds N.
    Returns (rptrG, cptrG), ascending columns, diagonal last."""
    nn = rptr.size - 1
    row_of = np.repeat(np.arange(nn, dtype=np.int64), np.diff(rptr))
    low = cptr <= row_of
    lr, lc = row_of[low], cptr[low].astype(np.int64)
    lens = np.bincount(lr, minlength=nn)
    if lens.max(initial=0) > N:
        if absval is None:
            raise ValueError("rows with more than %d lower entries need |A| for the pattern" % N)
        v = np.asarray(absval)[low]
        # MATLAB find() enumerates tril(A) column-major: sort by (col, row)
        o = np.lexsort((lr, lc))
        i, j, v = lr[o], lc[o], v[o]
        order = np.lexsort((-v, -i))              # stable: row desc, |v| desc
        rs = i[order]
        d = np.diff(np.concatenate([[rs[-1] + 1], rs]))
        starts = np.flatnonzero(d) + 1
        cand = np.unique((starts[None, :] - np.arange(N)[:, None]).ravel())
        cand = cand[cand > 0]
        keep = order[cand - 1]
        lr, lc = i[keep], j[keep]
    o = np.lexsort((lc, lr))
    lr, lc = lr[o], lc[o]
    rp = np.zeros(nn + 1, np.int64)
    np.add.at(rp, lr + 1, 1)
    return np.cumsum(rp).astype(np.int32), lc.astype(np.int32)


def transpose_pattern(rptr: np.ndarray, cptr: np.ndarray):
    """CSR of G^T and perm with GT_value[p] = G_value[perm[p]] (the stable
    column re-sort of dMesh.m:1255-1256)."""
    nn = rptr.size - 1
    rows = np.repeat(np.arange(nn, dtype=np.int64), np.diff(rptr))
    cols = cptr.astype(np.int64)
    perm = np.lexsort((rows, cols))
    rp = np.zeros(nn + 1, np.int64)
    np.add.at(rp, cols + 1, 1)
    return np.cumsum(rp).astype(np.int32), rows[perm].astype(np.int32), perm.astype(np.int32)


def locate_points(node: np.ndarray, elem: np.ndarray, pts: np.ndarray, tol: float = 1e-9):
    """Containing tetrahedron and barycentric weights of each point (-1 and
    NaN outside).  Candidates are filtered by bounding box; among containing
    tets the one with the largest minimum weight is taken."""
    pts = np.atleast_2d(np.asarray(pts, float))
    x = node[elem]                                  # NE x 4 x 3
    lo = x.min(axis=1) - 1e-9
    hi = x.max(axis=1) + 1e-9
    eid = np.full(len(pts), -1, np.int64)
    w = np.full((len(pts), 4), np.nan)
    for i, p in enumerate(pts):
        cand = np.flatnonzero(np.all((lo <= p) & (p <= hi), axis=1))
        if cand.size == 0:
            continue
        xc = x[cand]
        T = np.stack([xc[:, 1] - xc[:, 0], xc[:, 2] - xc[:, 0], xc[:, 3] - xc[:, 0]], axis=2)
        lam = np.linalg.solve(T, (p[None, :] - xc[:, 0])[:, :, None])[:, :, 0]
        ww = np.concatenate([1.0 - lam.sum(axis=1, keepdims=True), lam], axis=1)
        k = int(np.argmax(ww.min(axis=1)))
        if ww[k].min() >= -tol:
            eid[i] = cand[k]
            w[i] = ww[k]
    return eid, w


def closest_point_on_triangle(p, a, b, c):
    """Closest point of triangle abc to p (Voronoi-region case analysis)."""
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
        return a + ab * (d1 / (d1 - d3))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + ac * (d2 / (d2 - d6))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))
    s = 1.0 / (va + vb + vc)
    return a + ab * (vb * s) + ac * (vc * s)


def clean_bary(w):
    """Weights below 1e-10 set to 0, then renormalised (dOptode.m:93-95)."""
    w = np.array(w, float)
    w[w < 1e-10] = 0.0
    return w / w.sum(axis=1, keepdims=True)


def snap_to_surface(node, face, bnd, pos):
    """Nearest point on the boundary surface via the nearest boundary node's
    incident faces, and the inward normal of the face hit
    (dOptode.m:155-199)."""
    pos = np.array(np.atleast_2d(pos), float)
    bidx = np.flatnonzero(bnd)
    normals = np.zeros_like(pos)
    for i in range(len(pos)):
        nb = bidx[np.argmin(np.linalg.norm(node[bidx] - pos[i], axis=1))]
        fi = np.flatnonzero(np.any(face == nb, axis=1))
        best, bp, bf = np.inf, None, None
        for f in fi:
            t = node[face[f]]
            q = closest_point_on_triangle(pos[i], t[0], t[1], t[2])
            d = np.linalg.norm(q - pos[i])
            if d < best:
                best, bp, bf = d, q, f
        pos[i] = bp
        t = node[face[bf]]
        n = -np.cross(t[1] - t[0], t[2] - t[0])
        normals[i] = n / np.linalg.norm(n)
    return pos, normals


def _spread3(v: np.ndarray) -> np.ndarray:
    """Insert two zero bits between the low 21 bits of v (3-D Morton code)."""
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def morton_order(node: np.ndarray, elem: np.ndarray, cell: float | None = None) -> np.ndarray:
    """Node order along a 3-D Z-order (Morton) curve of the coordinates
    quantised to ``cell`` (default: the median edge length).  Consecutive
    runs of 8^k nodes then form compact bricks, so a tile of consecutive rows
    has a small stencil halo.  Returns order[i] = node placed at row i."""
    if cell is None:
        e = elem[: min(len(elem), 200000)]
        d = np.linalg.norm(node[e[:, 1]] - node[e[:, 0]], axis=1)
        cell = float(np.median(d[d > 0])) if np.any(d > 0) else 1.0
        # Kuhn and other structured tets have face/body diagonals among the
        # edges: use the shortest edge class so lattice points map 1:1
        cell = min(cell, float(np.percentile(d[d > 0], 5)))
    q = np.floor((node - node.min(axis=0)) / cell + 0.5).astype(np.int64)
    q = np.clip(q, 0, (1 << 21) - 1)
    code = _spread3(q[:, 0]) | (_spread3(q[:, 1]) << np.uint64(1)) | (_spread3(q[:, 2]) << np.uint64(2))
    return np.argsort(code, kind="stable").astype(np.int32)
