"""Pins the Jacobian oracle: the exact derivative against central finite
differences of meas_flu in nodal mua (kappa fixed), and the reference's
nodal-product approximation (J_complex) against the exact derivative."""
import numpy as np
import pytest

from oracle import fem, jacobian as ojac, optode as oopt, solver as osol, transforms as otr


@pytest.fixture(scope="module", params=[1.0, 2.0])
def problem(request):
    from dogpup_amd import meshgen
    h = request.param
    m = meshgen.slab((20.0, 20.0, 10.0), h, 0.01, 1.0, 1.4)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    t = np.arange(0, 501) * 20e-12
    f, df = otr.freq_axis(t, 3)
    A = fem.full_system(om, df)
    op = oopt.snap2mesh(om, np.array([[6.0, 10.0, 10.0]]), np.array([[14.0, 10.0, 10.0]]))
    sf = np.array([1.0, 0.8 - 0.2j, 0.5 + 0.1j])
    Q = oopt.source_matrix(om, op["s_id"], op["s_w"], sf)
    QA = oopt.source_matrix(om, op["d_id"], op["d_w"], sf)
    phi, phiA = osol.direct(A, Q), osol.direct(A, QA)
    link = np.array([[0, 0]])
    nodes = [int(np.argmin(np.linalg.norm(m.node - np.array(p), axis=1))) for p in [(10, 10, 4), (8, 10, 6)]]
    return dict(h=h, m=m, om=om, df=df, Q=Q, op=op, phi=phi, phiA=phiA, link=link, sf=sf, nodes=nodes)


def _data(P, mua):
    om = P["om"]
    om2 = fem.OMesh(node=om.node, elem=om.elem, nr=om.nr, mua=mua, musp=om.musp, kappa=om.kappa.copy())
    fem.prepare_mesh(om2)
    om2.K, om2.M = fem.gen_fwdmat(om2)
    ph = osol.direct(fem.full_system(om2, P["df"]), P["Q"])
    return oopt.meas_flu(om2, P["op"]["d_id"], P["op"]["d_w"], P["link"], ph)


def test_exact_derivative_vs_central_fd(problem):
    P = problem
    Je = ojac.j_exact(P["om"], P["phi"], P["phiA"], P["link"], P["sf"], P["nodes"])
    for q, v in enumerate(P["nodes"]):
        d = 1e-5
        mp, mm = P["om"].mua.copy(), P["om"].mua.copy()
        mp[v] += d
        mm[v] -= d
        fd = (_data(P, mp) - _data(P, mm)) / (2 * d)
        assert np.abs(fd[0] - Je[0, :, q]).max() < 1e-6 * np.abs(Je[0, :, q]).max()


def test_nodal_product_approximates_exact(problem):
    P = problem
    Jn = ojac.j_complex(P["phi"], P["phiA"], P["link"], P["sf"], P["h"])
    Je = ojac.j_exact(P["om"], P["phi"], P["phiA"], P["link"], P["sf"], P["nodes"])
    for q, v in enumerate(P["nodes"]):
        r = np.abs(Jn[0, :, v] - Je[0, :, q]) / np.abs(Je[0, :, q])
        assert r.max() < 0.15, r


def test_j_tg_is_fc2tg_of_j_complex(problem):
    P = problem
    t = np.arange(0, 501) * 20e-12
    f, _ = otr.freq_axis(t, 3)
    gates = np.stack([((t >= a) & (t < a + 500e-12)).astype(float) for a in (0.0, 2e-10)])
    G = otr.td2fc(gates, f, t, 1)
    J = ojac.j_complex(P["phi"], P["phiA"], P["link"], P["sf"], P["h"])
    Jtg = ojac.j_tg(P["phi"], P["phiA"], P["link"], P["sf"], P["h"], G, f, t.size)
    # explicit: J_TG(m,g,v) = Nt 2 Re sum_k conj(G_gk) w_k J(m,k,v)
    w = np.array([0.5, 1.0, 1.0])
    expect = t.size * 2 * np.real(np.einsum("gk,k,mkv->mgv", np.conj(G), w, J))
    assert np.allclose(Jtg, expect, rtol=1e-12, atol=0)
