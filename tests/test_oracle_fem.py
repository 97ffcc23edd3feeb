"""Pins the FEM oracle (oracle/fem.py) to closed forms and invariants."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import small_slab
from oracle import fem


def full(Kt):
    return (Kt + sp.triu(Kt, 1).T).tocsr()


def test_robin_constant():
    # n = 1.4: R0 = 0.0277778, A = 2.743860, R = 1/(2A) = 0.1822250
    assert fem.robin_factor(1.4) == pytest.approx(0.1822250, abs=5e-8)


def test_kuhn_counts_config1():
    from dogpup_amd import meshgen
    m, _, _, _ = meshgen.slab_60x60x30_h2()
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    assert om.node.shape[0] == 15376 and om.elem.shape[0] == 81000
    assert om.face.shape[0] == 7200 and int(om.bnd.sum()) == 3602
    assert fem.boundary_faces_ref(om).shape[0] - om.face.shape[0] == 1452
    A = fem.full_system(om, 1.0)
    assert A.nnz == 215326 and sp.tril(A).nnz == 115351
    assert om.vol.sum() == pytest.approx(108000.0, rel=1e-12)


def single_tet_mesh(pts, mua, kappa_musp=1.0):
    node = np.asarray(pts, float)
    elem = np.array([[0, 1, 2, 3]])
    return fem.make_mesh(node, elem, 1.4, mua, kappa_musp)


def test_element_matrices_closed_form():
    rng = np.random.default_rng(0)
    pts = np.array([[0, 0, 0], [1.3, 0.1, 0], [0.2, 0.9, 0.1], [0.1, 0.3, 1.1]]) + rng.normal(0, 0.05, (4, 3))
    mua = np.array([0.01, 0.02, 0.03, 0.05])
    m = single_tet_mesh(pts, mua, 1.0)
    K = full(m.K).toarray()
    M = full(m.M).toarray()
    vol = abs(np.linalg.det(np.stack([pts[1] - pts[0], pts[2] - pts[0], pts[3] - pts[0]]))) / 6
    # gradients of barycentric coordinates by solving the 4x4 interpolation system
    V = np.hstack([np.ones((4, 1)), pts])
    grads = np.linalg.inv(V)[1:, :].T
    kbar = m.kappa.mean()
    Kexp = vol * kbar * grads @ grads.T
    for a, b in itertools.product(range(4), range(4)):
        s = 0.0
        for k in range(4):
            c = sorted([[a, b, k].count(u) for u in set([a, b, k])], reverse=True)
            s += mua[k] * {(3,): 1 / 20, (2, 1): 1 / 60, (1, 1, 1): 1 / 120}[tuple(c)]
        Kexp[a, b] += vol * s
    # Robin: a single tet has all 4 nodes on the boundary -> 4 faces
    R = fem.robin_factor(1.4)
    for f in [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]:
        area = 0.5 * np.linalg.norm(np.cross(pts[f[1]] - pts[f[0]], pts[f[2]] - pts[f[0]]))
        for i, j in itertools.product(range(3), range(3)):
            Kexp[f[i], f[j]] += R * area * (1 + (i == j)) / 12
    Mexp = vol * (np.ones((4, 4)) + np.eye(4)) / 20
    assert np.allclose(K, Kexp, rtol=1e-12, atol=1e-15)
    assert np.allclose(M, Mexp, rtol=1e-13)


def test_assembly_invariants(tiny_mesh):
    m = tiny_mesh
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    K = full(om.K)
    M = full(om.M)
    one = np.ones(om.node.shape[0])
    assert one @ M @ one == pytest.approx(om.vol.sum(), rel=1e-12)
    # stiffness annihilates constants: K - absorption - Robin applied to 1
    om0 = fem.make_mesh(m.node, m.elem, m.nr, np.zeros_like(m.mua), m.musp)
    Kst = full(om0.K)
    bf = fem.boundary_faces_ref(om0)
    a = m.node[bf[:, 1]] - m.node[bf[:, 0]]
    b = m.node[bf[:, 2]] - m.node[bf[:, 0]]
    area = 0.5 * np.linalg.norm(np.cross(a, b), axis=1)
    R = fem.robin_factor(1.4)
    # 1'K1 = int mua + R * (sum of Robin-face areas)
    assert one @ Kst @ one == pytest.approx(R * area.sum(), rel=1e-11)
    mua_int = one @ (K - Kst) @ one
    assert mua_int == pytest.approx(0.01 * om.vol.sum(), rel=1e-11)
    # K is symmetric positive definite
    assert abs(K - K.T).max() < 1e-15 * abs(K).max()


def test_reference_face_list_slab():
    # The reference's Robin face list on a regular Kuhn slab: the true faces,
    # plus interior faces whose 3 nodes are all on the boundary (both
    # neighbouring tets contribute them).  60x60x30 mm, h = 2: 7,200 true
    # faces + 1,452 extra = 288 axis faces (area h^2/2) + 1,164 diagonal
    # faces (area h^2*sqrt(2)/2).
    from dogpup_amd import meshgen
    m = meshgen.slab((60.0, 60.0, 30.0), 2.0, 0.01, 1.0, 1.4)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    bf = fem.boundary_faces_ref(om)
    a = m.node[bf[:, 1]] - m.node[bf[:, 0]]
    b = m.node[bf[:, 2]] - m.node[bf[:, 0]]
    area = 0.5 * np.linalg.norm(np.cross(a, b), axis=1) / 4.0
    half = np.isclose(area, 0.5)
    diag = np.isclose(area, np.sqrt(2) / 2)
    assert half.sum() == 7200 + 288 and diag.sum() == 1164 and bf.shape[0] == 8652
    K = full(om.K)
    one = np.ones(om.node.shape[0])
    expect = 0.01 * 108000.0 + fem.robin_factor(1.4) * 4.0 * (0.5 * half.sum() + np.sqrt(2) / 2 * diag.sum())
    assert one @ K @ one == pytest.approx(expect, rel=1e-12)
