"""Pins the FSAI and BiCGSTAB oracle: exact FSAI identities, brute-force
per-row solves, and the converged BiCGSTAB solution vs a direct LU solve."""
import numpy as np
import scipy.sparse as sp

from conftest import rel, small_slab
from oracle import fem, fsai as ofsai, optode as oopt, solver as osol, transforms as otr


def system(seed=1):
    m = small_slab(seed=seed, inclusions=[((10.0, 10.0, 5.0), 3.0, 0.03)])
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    t = np.arange(0, 501) * 20e-12
    f, df = otr.freq_axis(t, 4)
    return m, om, fem.full_system(om, df), t, f


def test_fsai_identities():
    _, _, A, _, _ = system()
    rG, cG, valG = ofsai.fsaip_gen(A, 3)
    n = A.shape[0]
    for k in range(3):
        Ak = fem.system_at(A, k)
        G = ofsai.g_matrix(rG, cG, valG[:, k], n)
        assert np.abs((G @ Ak @ G.T).diagonal() - 1).max() < 1e-11
        GA = (G @ Ak).tocsr()
        rows = np.repeat(np.arange(n), np.diff(rG))
        off = cG != rows
        vals = np.asarray(GA[rows[off], cG[off]]).ravel()
        assert np.abs(vals).max() < 1e-13 * np.abs(GA.data).max()


def test_fsai_diagonal_matrix():
    d = np.array([4.0, 9.0 + 2j, 2.0 - 1j, 16.0])
    A = sp.diags(d).tocsr()
    rG, cG, valG = ofsai.fsaip_gen(sp.csr_matrix(A.real) + 1j * sp.csr_matrix(A.imag), 2)
    assert (cG == np.arange(4)).all()
    assert np.allclose(valG[:, 1], 1 / np.sqrt(d))
    assert np.allclose(valG[:, 0], 1 / np.sqrt(d.real))


def test_fsai_rows_brute_force_and_batched():
    _, _, A, _, _ = system(2)
    rG, cG, valG = ofsai.fsaip_gen(A, 2)
    r2, c2, v2 = ofsai.fsaip_gen_batched(A, 2)
    assert (rG == r2).all() and (cG == c2).all() and rel(v2, valG) < 1e-12
    # brute force: g = x / sqrt(x_m), A[P,P] x = e_m with numpy.linalg.solve
    Ak = fem.system_at(A, 1).toarray()
    for i in [0, 17, 100, A.shape[0] - 1]:
        P = cG[rG[i]:rG[i + 1]]
        e = np.zeros(len(P))
        e[-1] = 1
        x = np.linalg.solve(Ak[np.ix_(P, P)], e)
        assert np.allclose(valG[rG[i]:rG[i + 1], 1], x / np.sqrt(x[-1]), rtol=1e-12)


def test_bicgstab_matches_direct():
    m, om, A, t, f = system(3)
    src = np.array([[5.0, 5.0, 10.0], [15.0, 12.0, 10.0]])
    det = np.array([[8.0, 8.0, 10.0]])
    op = oopt.snap2mesh(om, src, det)
    sf = np.exp(-1j * f * 1e-9) * (1 + 0.1 * np.arange(f.size))
    Q = oopt.source_matrix(om, op["s_id"], op["s_w"], sf)
    x, it, err = osol.bicgstab_fsai(A, Q, tol=1e-12)
    xd = osol.direct(A, Q)
    assert err < 1e-12 and it < 100
    assert rel(x, xd) < 1e-10
    # residual of the direct solve
    for k in range(Q.shape[1]):
        Ak = fem.system_at(A, k)
        assert rel(Ak @ xd[:, k, :], Q[:, k, :]) < 1e-12
