"""Synthetic slab meshes and optode layouts for the frequency-domain diffusion path.

The reference builds its meshes with iso2mesh; the Kuhn split used here has the
behaviour of ``DOGPUP/meshing/iso2mesh/meshgrid6.m:35-67``: every lattice cube is
cut into the 6 tetrahedra that share its main diagonal, and the lattice is
numbered with x fastest (MATLAB ``ndgrid`` order).  Node numbering is 0-based
here; element orientation is made positive (as ``meshreorient`` does).

Everything is seeded so the oracle and the HIP path see identical inputs.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np


@dataclass
class SlabMesh:
    """A Kuhn-split box mesh plus its nodal optical properties."""

    node: np.ndarray          # NN x 3 float64 (mm)
    elem: np.ndarray          # NE x 4 int32, 0-based
    mua: np.ndarray           # NN float64 (1/mm)
    musp: np.ndarray          # NN float64 (1/mm)
    nr: float                 # refractive index
    shape: tuple              # (nx, ny, nz) lattice points
    h: float                  # lattice spacing (mm)
    extent: tuple             # (Lx, Ly, Lz) box size (mm)
    meta: dict = field(default_factory=dict)


def kuhn_lattice(nx: int, ny: int, nz: int, h: float):
    """Regular lattice nodes and the 6-tet Kuhn split of every cube.

    Tetrahedron for a permutation (p1, p2, p3) of the axes holds the cube
    corners v with v[p1] <= v[p2] <= v[p3] (``meshgrid6.m:35-42``).
    """
    xs = np.arange(nx) * h
    ys = np.arange(ny) * h
    zs = np.arange(nz) * h
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    node = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)  # x fastest
    stride = np.array([1, nx, nx * ny], dtype=np.int64)

    ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    base = (ci.ravel(order="F") * stride[0] + cj.ravel(order="F") * stride[1]
            + ck.ravel(order="F") * stride[2]).astype(np.int64)
    # cube corner offsets (bit b of corner id -> axis b)
    corners = np.array([[(c >> a) & 1 for a in range(3)] for c in range(8)], dtype=np.int64)
    tets = []
    for perm in itertools.permutations(range(3)):
        verts = [c for c in range(8)
                 if corners[c][perm[0]] <= corners[c][perm[1]] <= corners[c][perm[2]]]
        assert len(verts) == 4
        off = corners[verts] @ stride
        tets.append(base[:, None] + off[None, :])
    elem = np.concatenate(tets, axis=0)
    return node, elem


def orient_positive(node: np.ndarray, elem: np.ndarray) -> np.ndarray:
    """Swap the last two vertices of negatively oriented tetrahedra."""
    a = node[elem[:, 1]] - node[elem[:, 0]]
    b = node[elem[:, 2]] - node[elem[:, 0]]
    c = node[elem[:, 3]] - node[elem[:, 0]]
    det = np.einsum("ij,ij->i", np.cross(a, b), c)
    elem = elem.copy()
    neg = det < 0
    elem[neg, 2], elem[neg, 3] = elem[neg, 3].copy(), elem[neg, 2].copy()
    return elem


def slab(extent=(60.0, 60.0, 30.0), h=2.0, mua=0.01, musp=1.0, nr=1.4,
         jitter=0.0, seed=0, inclusions=()):
    """Kuhn slab ``extent`` mm with spacing ``h``.

    ``jitter`` moves interior nodes by U(-jitter, jitter) mm per axis.
    ``inclusions`` is a list of (centre xyz, radius, mua) spheres; nodes inside
    take the inclusion's absorption.
    """
    nx, ny, nz = (int(round(e / h)) + 1 for e in extent)
    node, elem = kuhn_lattice(nx, ny, nz, h)
    rng = np.random.default_rng(seed)
    if jitter > 0:
        ijk = np.stack(np.unravel_index(np.arange(node.shape[0]), (nz, ny, nx))[::-1], axis=1)
        interior = np.all((ijk > 0) & (ijk < np.array([nx - 1, ny - 1, nz - 1])), axis=1)
        node = node.copy()
        node[interior] += rng.uniform(-jitter, jitter, size=(int(interior.sum()), 3))
    elem = orient_positive(node, elem).astype(np.int32)
    mua_n = np.full(node.shape[0], float(mua))
    musp_n = np.full(node.shape[0], float(musp))
    for centre, radius, mua_inc in inclusions:
        inside = np.linalg.norm(node - np.asarray(centre, float)[None, :], axis=1) <= radius
        mua_n[inside] = mua_inc
    return SlabMesh(node=node, elem=elem, mua=mua_n, musp=musp_n, nr=float(nr),
                    shape=(nx, ny, nz), h=float(h), extent=tuple(float(e) for e in extent),
                    meta={"seed": seed, "jitter": jitter, "inclusions": list(inclusions)})


def top_grid(extent, nx, ny, x0, y0, dx, dy):
    """Points on the top face z = Lz on an nx-by-ny grid (x fastest)."""
    xs = x0 + dx * np.arange(nx)
    ys = y0 + dy * np.arange(ny)
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, extent[2])], axis=1)


def all_pairs(ns: int, nd: int) -> np.ndarray:
    """Link table [source, detector] (0-based) for every pair, source fastest."""
    s, d = np.meshgrid(np.arange(ns), np.arange(nd), indexing="xy")
    return np.stack([s.ravel(), d.ravel()], axis=1).astype(np.int32)


# --- named workloads ------------------------------------------------------------

def slab_60x60x30_h2(seed=0):
    """60x60x30 mm slab, 2 mm Kuhn grid (15,376 nodes), interior jitter 0.2 mm,
    mua 0.01/mm, musp 1.0/mm, n 1.4; 16 sources and 16 detectors on 4x4 grids
    of the top face."""
    m = slab((60.0, 60.0, 30.0), 2.0, 0.01, 1.0, 1.4, jitter=0.2, seed=seed)
    src = top_grid(m.extent, 4, 4, 12.0, 12.0, 12.0, 12.0)
    det = top_grid(m.extent, 4, 4, 18.0, 18.0, 12.0, 12.0)
    return m, src, det, all_pairs(16, 16)


def sorted_mesh(m: SlabMesh) -> SlabMesh:
    """The same mesh with its nodes renumbered along the solver's Morton
    curve and its elements sorted (meshio.sortmesh, the purpose of
    iso2mesh's sortmesh); nodal properties follow their nodes.  A solver
    handle made on it works in the caller's node order directly: the row
    permutations of b and x in and out of every solve are skipped."""
    from .meshio import sortmesh
    node, elem, order = sortmesh(m.node, m.elem)
    meta = dict(m.meta)
    meta["morton_sorted"] = True
    return SlabMesh(node, elem.astype(np.int32), m.mua[order], m.musp[order], m.nr, m.shape, m.h, m.extent, meta)


def slab_80x80x40_h1(seed=0, inclusion=True):
    """80x80x40 mm slab, 1 mm Kuhn grid (269,001 nodes, 1,536,000 tets),
    mua 0.01/mm, musp 1.0/mm, n 1.4, one 6 mm sphere of mua 0.03/mm at a seeded
    depth of 10-25 mm below the top face; 32 sources and 32 detectors on two
    interleaved 4x8 grids of the top face."""
    rng = np.random.default_rng(seed)
    depth = rng.uniform(10.0, 25.0)
    inc = [((40.0, 40.0, 40.0 - depth), 6.0, 0.03)] if inclusion else []
    m = slab((80.0, 80.0, 40.0), 1.0, 0.01, 1.0, 1.4, jitter=0.0, seed=seed, inclusions=inc)
    src = top_grid(m.extent, 4, 8, 16.0, 5.0, 16.0, 10.0)
    det = top_grid(m.extent, 4, 8, 24.0, 10.0, 16.0, 10.0)
    return m, src, det, all_pairs(32, 32)


def gaussian_tpsf(t, fwhm=250e-12, centre_sigmas=6.0):
    """Gaussian source pulse sampled on ``t`` (s)."""
    sd = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    return np.exp(-((t - centre_sigmas * sd) ** 2) / (2.0 * sd ** 2))


def boxcar_gates(t, n_gates=8, width=500e-12, spacing=200e-12, start=0.0):
    """n_gates boxcar time gates of ``width`` s starting ``spacing`` s apart."""
    g = np.zeros((n_gates, t.size))
    for i in range(n_gates):
        t0 = start + i * spacing
        g[i] = ((t >= t0) & (t < t0 + width)).astype(float)
    return g


def cylinder(radius, height, h=2.0, mua=0.01, musp=1.0, nr=1.4, seed=0):
    """Cylinder of ``radius`` and ``height`` mm (axis z, centre (radius,
    radius)): the Kuhn lattice of its bounding box keeps the tetrahedra whose
    four nodes lie inside the radius (a staircase side wall, flat top and
    bottom); unused nodes are dropped and the rest renumbered in lattice
    order."""
    n = int(round(2.0 * radius / h)) + 1
    nz = int(round(height / h)) + 1
    node, elem = kuhn_lattice(n, n, nz, h)
    c = (n - 1) * h / 2.0
    inside = (node[:, 0] - c) ** 2 + (node[:, 1] - c) ** 2 <= radius ** 2 * (1.0 + 1e-9)
    elem = elem[inside[elem].all(axis=1)]
    used = np.unique(elem)
    remap = np.full(node.shape[0], -1, np.int64)
    remap[used] = np.arange(used.size)
    node = node[used]
    elem = orient_positive(node, remap[elem]).astype(np.int32)
    return SlabMesh(node=node, elem=elem, mua=np.full(node.shape[0], float(mua)),
                    musp=np.full(node.shape[0], float(musp)), nr=float(nr), shape=(n, n, nz), h=float(h),
                    extent=(2.0 * c, 2.0 * c, (nz - 1) * h), meta={"seed": seed, "kind": "cylinder",
                                                                 "radius": radius})


def uneven_batch(n=48, seed=0, nn_range=(5000, 200000)):
    """Seeded batch of slab and cylinder meshes of uneven size: node counts
    log-uniform in ``nn_range``, lattice spacing 1, 2 or 3 mm, mua U[0.005,
    0.02]/mm and musp U[0.5, 1.5]/mm, 8-32 sources on the top face, 8-64
    frequencies of 50 MHz spacing (T = 20 ns).  Returns a list of dicts with
    ``mesh``, ``src``, ``det`` (4 detectors), ``link`` and ``nf``."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        target = float(np.exp(rng.uniform(np.log(nn_range[0]), np.log(nn_range[1]))))
        h = float(rng.choice([1.0, 2.0, 3.0]))
        mua = float(rng.uniform(0.005, 0.02))
        musp = float(rng.uniform(0.5, 1.5))
        if k % 2 == 0:
            # slab with aspect 1 : 0.8 : 0.5
            nx = max(int(round((target / 0.4) ** (1.0 / 3.0))), 6)
            ny, nz = max(int(round(0.8 * nx)), 5), max(int(round(0.5 * nx)), 4)
            m = slab(((nx - 1) * h, (ny - 1) * h, (nz - 1) * h), h, mua, musp, 1.4, seed=seed * 1000 + k)
        else:
            # cylinder of height = radius, about target nodes (pi/4 of its box)
            nb = max(int(round((target / (np.pi / 4.0) * 2.0) ** (1.0 / 3.0))), 8)
            m = cylinder((nb - 1) * h / 2.0, (nb // 2) * h, h, mua, musp, 1.4, seed=seed * 1000 + k)
        ns = int(rng.integers(8, 33))
        nf = int(rng.integers(8, 65))
        gx = int(np.ceil(np.sqrt(ns)))
        gy = -(-ns // gx)
        ext = m.extent
        src = top_grid(ext, gx, gy, 0.2 * ext[0], 0.2 * ext[1], 0.6 * ext[0] / max(gx - 1, 1),
                       0.6 * ext[1] / max(gy - 1, 1))[:ns]
        det = top_grid(ext, 2, 2, 0.3 * ext[0], 0.3 * ext[1], 0.4 * ext[0], 0.4 * ext[1])
        out.append({"mesh": m, "src": src, "det": det, "link": all_pairs(ns, 4), "nf": nf})
    return out


def batch_48_uneven(seed=1):
    """The 48-instance uneven batch: seed 1 draws 2.59M nodes in total (mean
    54k), 38,431 RHS and 1.96e9 node x RHS."""
    return uneven_batch(48, seed)


def slab_80x80x40_two_inclusions(h=1.0, seed=0):
    """80x80x40 mm slab on an h mm Kuhn grid, background mua 0.01/mm and
    musp 1.0/mm, with two absorbing spheres (radius 5 mm, mua 0.02/mm and
    radius 7 mm, mua 0.025/mm) at seeded depths of 10-20 mm, 24 mm apart; 32
    sources and 32 detectors as in slab_80x80x40_h1.  Returns (mesh, src, det,
    link, inclusions)."""
    rng = np.random.default_rng(seed)
    d1, d2 = rng.uniform(10.0, 20.0, size=2)
    inc = [((28.0, 40.0, 40.0 - d1), 5.0, 0.02), ((52.0, 40.0, 40.0 - d2), 7.0, 0.025)]
    m = slab((80.0, 80.0, 40.0), h, 0.01, 1.0, 1.4, jitter=0.0, seed=seed, inclusions=inc)
    src = top_grid(m.extent, 4, 8, 16.0, 5.0, 16.0, 10.0)
    det = top_grid(m.extent, 4, 8, 24.0, 10.0, 16.0, 10.0)
    return m, src, det, all_pairs(32, 32), inc
