"""Oracle voxel basis (mesh2grid, dMesh.m:286-380) and mesh2mesh (:383-411).

Pinned by what fixes them independently of the restatement:
  - the grid's Kuhn split is a Delaunay triangulation of the grid points (no
    grid point inside any tetrahedron's circumsphere) and tiles every cell;
  - m2g and g2m reproduce linear functions exactly wherever all weights are
    kept (every triangulation does), rows of g2m sum to 1, and on a lattice
    mesh whose nodes are the grid points both are identities;
  - hand-computed rows for the in-mesh filter, the renormalisation, the
    zero-sum quirk and the nearest-voxel fallback.
"""

import numpy as np
import pytest

from conftest import small_slab
from oracle import grid as og


def test_kuhn_split_is_delaunay_and_tiles_cells():
    x, y, z = np.arange(3) * 1.0, np.arange(4) * 1.5, np.arange(3) * 2.0
    g = og.ndgrid_points(x, y, z)
    tets = og.kuhn_tets(len(x), len(y), len(z))
    assert tets.shape[0] == 6 * 2 * 3 * 2
    P = g[tets]                                                    # T x 4 x 3
    vol = np.abs(np.einsum("ti,ti->t", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0])) / 6
    assert np.allclose(vol, 1.0 * 1.5 * 2.0 / 6)                   # six equal tets per cell
    for t in range(P.shape[0]):
        A = 2 * (P[t, 1:] - P[t, 0])
        b = (P[t, 1:] ** 2).sum(1) - (P[t, 0] ** 2).sum()
        c = np.linalg.solve(A, b)
        r2 = ((P[t, 0] - c) ** 2).sum()
        d2 = ((g - c) ** 2).sum(1)
        assert (d2 >= r2 * (1 - 1e-12)).all()                      # empty circumsphere


def _lin(p):
    return 0.3 + 1.7 * p[:, 0] - 0.4 * p[:, 1] + 2.2 * p[:, 2]


def test_mesh2grid_linear_reproduction_and_inside_flags():
    m = small_slab(seed=4, extent=(12.0, 10.0, 8.0), h=2.0)
    x = -1.3 + 1.7 * np.arange(9)
    y = -0.9 + 1.3 * np.arange(10)
    z = -0.7 + 1.1 * np.arange(9)
    r = og.mesh2grid(m.node, m.elem, x, y, z)
    g = r["grid"]
    inside = np.all((g >= 0) & (g <= np.array([12.0, 10.0, 8.0])), axis=1)
    assert np.array_equal(r["gridinmesh"], inside)
    f = _lin(m.node)
    got = r["m2g"] @ f
    assert np.allclose(got[inside], _lin(g[inside]), atol=1e-12)
    assert np.all(got[~inside] == 0)
    # g2m: rows sum to 1; exact where a node's grid tetrahedron is all in-mesh
    s = np.asarray(r["g2m"].sum(axis=1)).ravel()
    assert np.allclose(s, 1.0)
    full = np.asarray((r["g2m"] != 0).sum(axis=1)).ravel() == 4
    gv = r["g2m"] @ _lin(g)
    assert full.sum() > 10
    assert np.allclose(gv[full], f[full], atol=1e-12)


def test_lattice_grid_gives_identities():
    m = small_slab(seed=2, jitter=0.0, extent=(6.0, 4.0, 4.0), h=2.0)
    x, y, z = np.arange(0, 7, 2.0), np.arange(0, 5, 2.0), np.arange(0, 5, 2.0)
    r = og.mesh2grid(m.node, m.elem, x, y, z)
    assert r["gridinmesh"].all()
    order = np.lexsort((m.node[:, 0], m.node[:, 1], m.node[:, 2]))   # x fastest, like the grid
    P = np.eye(m.node.shape[0])[order]                              # grid v <- node order[v]
    assert np.allclose(r["m2g"].toarray(), P, atol=1e-12)
    assert np.allclose(r["g2m"].toarray(), P.T, atol=1e-12)
    assert r["dxyz"] == 2.0 and r["gridSize"] == (4, 3, 3)


@pytest.mark.parametrize("inmesh_ids,expect", [
    ([6], {6: 1.0}),                              # no vertex of its tetrahedron in mesh -> nearest voxel
    ([1, 6], {1: 1.0}),                           # one kept vertex renormalised
    ([1, 3], {1: 2.0 / 3.0, 3: 1.0 / 3.0}),       # 0.1 and 0.05 renormalised
    ([0, 1, 3, 7], {0: 0.8, 1: 0.1, 3: 0.05, 7: 0.05}),
])
def test_grid_to_mesh_hand_rows(inmesh_ids, expect):
    # unit cube grid; node f = (0.2, 0.1, 0.05): tetrahedron 0, e_x, e_x+e_y,
    # (1,1,1) = voxels 0, 1, 3, 7 with weights 0.8, 0.1, 0.05, 0.05
    x = y = z = np.array([0.0, 1.0])
    inm = np.zeros(8, bool)
    inm[inmesh_ids] = True
    g2m, fb = og.grid_to_mesh(np.array([[0.2, 0.1, 0.05]]), x, y, z, inm)
    row = g2m.toarray()[0]
    want = np.zeros(8)
    for k, v in expect.items():
        want[k] = v
    assert np.allclose(row, want, atol=1e-15)
    assert (fb.size == 1) == (len(expect) == 1 and 6 in expect)


def test_grid_to_mesh_zero_sum_row_stays_zero():
    # node on voxel 0 (weight 1) with voxel 0 outside and voxel 1 (weight 0)
    # inside: the row keeps one zero entry, its sum 0 is left unscaled.
    x = y = z = np.array([0.0, 1.0])
    inm = np.zeros(8, bool)
    inm[1] = True
    g2m, fb = og.grid_to_mesh(np.array([[0.0, 0.0, 0.0]]), x, y, z, inm)
    assert fb.size == 0 and not g2m.toarray().any()


def test_mesh2mesh_linear_and_outside_fill():
    old = small_slab(seed=5, extent=(8.0, 8.0, 4.0), h=2.0)
    new = small_slab(seed=6, extent=(8.0, 8.0, 4.0), h=1.0)
    f = _lin(old.node)
    got = og.mesh2mesh(new.node, f, old.node, old.elem)[:, 0]
    assert np.allclose(got, _lin(new.node), atol=1e-12)
    pts = np.array([[1.0, 1.0, 1.0], [9.0, 1.0, 1.0]])             # second point outside
    got = og.mesh2mesh(pts, f, old.node, old.elem)[:, 0]
    assert np.allclose(got, _lin(pts[:1])[0])
