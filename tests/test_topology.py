"""Host-side topology of the product (dogpup_amd/topology.py) against the oracle."""
import numpy as np
import scipy.sparse as sp

from conftest import small_slab
from dogpup_amd import topology
from oracle import fem, fsai as ofsai


def emulate_slots(pat, node, elem, bface, mua, kappa, R):
    """CPU replay of the assembly's slot sums from closed-form element
    matrices (independent of the oracle's quadrature loops)."""
    ne = elem.shape[0]
    x = node[elem]
    B = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=2)
    Binv = np.linalg.inv(B)
    gr = np.concatenate([-Binv.sum(axis=1, keepdims=True), Binv], axis=1)    # NE x 4 x 3
    vol = np.abs(np.linalg.det(B)) / 6
    kb = kappa[elem].mean(axis=1)
    mu = mua[elem]
    EK = np.einsum("e,eak,ebk->eab", vol * kb, gr, gr)
    s = mu.sum(1)
    for a in range(4):
        for b in range(4):
            if a == b:
                EK[:, a, b] += vol * (mu[:, a] / 20 + (s - mu[:, a]) / 60)
            else:
                EK[:, a, b] += vol * ((mu[:, a] + mu[:, b]) / 60 + (s - mu[:, a] - mu[:, b]) / 120)
    EM = vol[:, None, None] * (np.ones((4, 4)) + np.eye(4))[None] / 20
    fa = node[bface[:, 1]] - node[bface[:, 0]]
    fb = node[bface[:, 2]] - node[bface[:, 0]]
    area = 0.5 * np.linalg.norm(np.cross(fa, fb), axis=1)
    EB = (R[bface].mean(axis=1) * area)[:, None, None] * (np.ones((3, 3)) + np.eye(3))[None] / 12
    valsK = np.concatenate([EK.reshape(-1), EB.reshape(-1)])
    valsM = np.concatenate([EM.reshape(-1), np.zeros(EB.size)])
    re = np.add.reduceat(valsK[pat.slot_src], pat.slot_ptr[:-1])
    mm = np.add.reduceat(valsM[pat.slot_src], pat.slot_ptr[:-1])
    return re, mm


def test_pattern_and_slots_match_oracle():
    m = small_slab(seed=5)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    face, area = topology.true_faces(m.node, m.elem)
    assert face.shape[0] == om.face.shape[0]
    bnd = np.zeros(m.node.shape[0], np.int32)
    bnd[np.unique(face.ravel())] = 1
    assert (bnd == om.bnd).all()
    bface = topology.robin_faces(m.elem, bnd)
    assert bface.shape == fem.boundary_faces_ref(om).shape
    pat = topology.system_pattern(m.node.shape[0], m.elem, bface)
    A = fem.full_system(om, 1.0)
    assert (pat.rptr == A.indptr).all() and (pat.cptr == A.indices).all()
    kappa = 1 / (3 * (m.mua + m.musp))
    R = topology.robin_factor(m.nr) * bnd
    re, mm = emulate_slots(pat, m.node, m.elem, bface, m.mua, kappa, R)
    K = (om.K + sp.triu(om.K, 1).T).tocsr()
    M = (om.M + sp.triu(om.M, 1).T).tocsr()
    K.sort_indices()
    M.sort_indices()
    assert np.abs(re - K.data).max() < 1e-13 * np.abs(K.data).max()
    assert np.abs(mm - M.data).max() < 1e-13 * np.abs(M.data).max()


def test_fsai_pattern_matches_oracle_rule():
    m = small_slab(seed=6)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    A = fem.full_system(om, 1.0)
    rG, cG = topology.fsai_pattern(A.indptr.astype(np.int32), A.indices.astype(np.int32))
    pr, pc = ofsai.pattern_rows_cols(sp.tril(A), 30)
    r2, c2 = ofsai.csr_from_pairs(pr, pc, A.shape[0])
    assert (rG == r2).all() and (cG == c2).all()
    # diagonal last in every row
    rows = np.repeat(np.arange(A.shape[0]), np.diff(rG))
    last = rG[1:] - 1
    assert (cG[last] == np.arange(A.shape[0])).all()


def test_fsai_pattern_long_rows_rule():
    """Rows with more than N lower entries: the reference's window keeps the
    largest entry and the N-1 smallest; checked on random matrices against
    the oracle's restatement with a small N."""
    rng = np.random.default_rng(2)
    for trial in range(20):
        n = 12
        M = sp.random(n, n, density=0.6, random_state=rng.integers(1 << 30)).tocsr()
        M = M + M.T + sp.eye(n) * 5
        M = sp.csr_matrix(M)
        M.sort_indices()
        for N in (2, 3, 4):
            rG, cG = topology.fsai_pattern(M.indptr.astype(np.int32), M.indices.astype(np.int32),
                                           np.abs(M.data), N=N)
            pr, pc = ofsai.pattern_rows_cols(sp.tril(M), N)
            r2, c2 = ofsai.csr_from_pairs(pr, pc, n)
            assert (rG == r2).all() and (cG == c2).all()


def test_transpose_pattern():
    m = small_slab(seed=7)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    A = fem.full_system(om, 1.0)
    rG, cG = topology.fsai_pattern(A.indptr.astype(np.int32), A.indices.astype(np.int32))
    vals = np.arange(cG.size, dtype=float) + 1
    rT, cT, perm = topology.transpose_pattern(rG, cG)
    G = sp.csr_matrix((vals, cG, rG), shape=A.shape)
    GT = sp.csr_matrix((vals[perm], cT, rT), shape=A.shape)
    assert abs(GT - G.T).max() == 0


def test_locate_and_snap():
    m = small_slab(seed=8)
    face, _ = topology.true_faces(m.node, m.elem)
    bnd = np.zeros(m.node.shape[0], np.int32)
    bnd[np.unique(face.ravel())] = 1
    pts = np.array([[5.3, 7.1, 10.4], [12.0, 3.3, 11.0]])
    pos, nrm = topology.snap_to_surface(m.node, face, bnd, pts)
    assert np.allclose(pos[:, 2], 10.0) and np.allclose(pos[:, :2], pts[:, :2])
    assert np.allclose(nrm, [[0, 0, -1], [0, 0, -1]])
    eid, w = topology.locate_points(m.node, m.elem, pos - np.array([0, 0, 1.0]))
    assert (eid >= 0).all()
    rec = np.einsum("pa,pad->pd", w, m.node[m.elem[eid]])
    assert np.allclose(rec, pos - np.array([0, 0, 1.0]))


def test_true_face_switch_covers_the_box_surface():
    """robin_faces without the reference's selection returns the true
    boundary faces: their area is the box surface; the reference's selection
    (dMesh.m:1129-1140) is a superset with extra interior faces."""
    m = small_slab(seed=6)
    face, area = topology.true_faces(m.node, m.elem)
    bnd = np.zeros(m.node.shape[0], np.int32)
    bnd[np.unique(face.ravel())] = 1
    tf = topology.robin_faces(m.elem, bnd, False, face)
    assert np.array_equal(tf, face)
    ex, ey, ez = m.extent
    assert abs(area.sum() - 2 * (ex * ey + ey * ez + ex * ez)) < 1e-9
    rf = topology.robin_faces(m.elem, bnd, True, face)
    key = lambda f: set(map(tuple, np.sort(f, axis=1).tolist()))
    assert key(face) < key(rf)
