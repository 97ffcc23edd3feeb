"""Independent pins of oracle functions whose form the product shares:
the FSAI pattern window (dMesh.m:1361-1375) against a literal brute force,
snap_pos (dOptode.m:77-97,155-199) against analytic placements on a lattice
mesh, and the Jacobian's link indexing (dMesh.m:239-245) against central
differences and reciprocity on several source/detector pairs."""
import numpy as np
import pytest
import scipy.sparse as sp


def _brute_window(A_tril, N):
    """The window of dMesh.m:1361-1375 as list operations: the entries of
    tril(A) listed row by row from the last row to the first, each row's
    entries by decreasing |a| (equal |a| in the column order MATLAB's find
    enumerates, a stable sort); kept: the first entry of every row's run and
    the N-1 entries listed just before it."""
    A = sp.csr_matrix(A_tril)
    A.sort_indices()
    listing, starts = [], []
    for i in range(A.shape[0] - 1, -1, -1):
        cols = A.indices[A.indptr[i]:A.indptr[i + 1]]
        vals = np.abs(A.data[A.indptr[i]:A.indptr[i + 1]])
        run = sorted(zip(cols.tolist(), vals.tolist()), key=lambda cv: -cv[1])   # stable: ties keep column order
        if run:
            starts.append(len(listing))
            listing += [(i, c) for c, _ in run]
    keep = set()
    for s in starts:
        for p in range(max(0, s - (N - 1)), s + 1):
            keep.add(listing[p])
    return sorted(keep)


def test_fsai_window_rule_brute_force():
    from dogpup_amd import topology
    from oracle import fsai as ofsai
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = int(rng.integers(6, 20))
        M = sp.random(n, n, density=float(rng.uniform(0.2, 0.9)), random_state=int(rng.integers(1 << 30))).tocsr()
        M = sp.csr_matrix(M + M.T + sp.eye(n) * 3.0)
        M.sort_indices()
        T = sp.tril(M).tocsr()
        for N in (2, 3, 5, 30):
            pr, pc = ofsai.pattern_rows_cols(T, N)
            got = sorted(zip(pr.tolist(), pc.tolist()))
            assert got == _brute_window(T, N), (trial, N)
            # the product's pattern is the same set
            rG, cG = topology.fsai_pattern(M.indptr.astype(np.int32), M.indices.astype(np.int32), np.abs(M.data), N=N)
            rows = np.repeat(np.arange(n), np.diff(rG))
            assert sorted(zip(rows.tolist(), cG.tolist())) == got
            # what the rule means: a row with more than N entries (every row
            # but row 0 has a lower neighbour listed after it) keeps its
            # largest entry and its N-1 smallest; a row of at most N keeps all
            for i in range(1, n):
                cols = T.indices[T.indptr[i]:T.indptr[i + 1]]
                vals = np.abs(T.data[T.indptr[i]:T.indptr[i + 1]])
                kept = {c for r, c in got if r == i}
                if len(cols) <= N:
                    assert kept == set(cols.tolist())
                else:
                    o = np.argsort(-vals, kind="stable")
                    assert kept == {int(cols[o[0]])} | set(cols[o[-(N - 1):]].tolist())


@pytest.fixture(scope="module")
def lattice():
    from dogpup_amd import meshgen
    from oracle import fem
    m = meshgen.slab((10.0, 8.0, 5.0), 1.0, 0.01, 1.0, 1.4)
    return m, fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)


def _node(m, p):
    return int(np.argmin(np.linalg.norm(m.node - np.asarray(p, float), axis=1)))


def test_snap_pos_on_lattice_is_analytic(lattice):
    """Optodes off a lattice surface land on the closest surface point with
    the nodal weights a P1 interpolant must have there: 1 at a node, 1/2 and
    1/2 at an edge midpoint; sources then move 1/min(musp) = 1 mm along the
    inward normal (dOptode.m:89), onto the lattice node below."""
    from oracle import optode as oopt
    m, om = lattice
    pts = np.array([[3.0, 4.0, 5.4],        # above a top-face node
                    [3.5, 4.0, 5.2],        # above a top-face edge midpoint
                    [-0.3, 2.0, 2.0],       # off the x = 0 face, level with a node
                    [6.0, 8.6, 3.0]])       # off the y = 8 face
    expect_pos = np.array([[3.0, 4.0, 5.0], [3.5, 4.0, 5.0], [0.0, 2.0, 2.0], [6.0, 8.0, 3.0]])
    pos, eid, w, nrm = oopt.snap_pos(om, pts)
    assert np.abs(pos - expect_pos).max() < 1e-12
    W = oopt.source_matrix(om, eid, w, np.ones(1))[:, 0, :]           # NN x 4 nodal weights
    for q, p in enumerate(expect_pos):
        if q == 1:
            a, b = _node(m, (3.0, 4.0, 5.0)), _node(m, (4.0, 4.0, 5.0))
            assert abs(W[a, q] - 0.5) < 1e-12 and abs(W[b, q] - 0.5) < 1e-12
        else:
            assert abs(W[_node(m, p), q] - 1.0) < 1e-12
        assert abs(W[:, q].sum() - 1.0) < 1e-12
    inward = np.array([[0, 0, -1], [0, 0, -1], [1, 0, 0], [0, -1, 0]], float)
    assert np.abs(nrm - inward).max() < 1e-12
    sn = oopt.snap2mesh(om, pts[[0, 2]], pts[[1]])
    assert np.abs(sn["s_pos"] - np.array([[3.0, 4.0, 4.0], [1.0, 2.0, 2.0]])).max() < 1e-12
    Ws = oopt.source_matrix(om, sn["s_id"], sn["s_w"], np.ones(1))[:, 0, :]
    assert abs(Ws[_node(m, (3.0, 4.0, 4.0)), 0] - 1.0) < 1e-12
    assert abs(Ws[_node(m, (1.0, 2.0, 2.0)), 1] - 1.0) < 1e-12


@pytest.fixture(scope="module")
def two_by_two():
    from dogpup_amd import meshgen
    from oracle import fem, optode as oopt, solver as osol, transforms as otr
    m = meshgen.slab((20.0, 16.0, 10.0), 1.0, 0.01, 1.0, 1.4)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    t = np.arange(0, 501) * 20e-12
    f, df = otr.freq_axis(t, 3)
    A = fem.full_system(om, df)
    op = oopt.snap2mesh(om, np.array([[5.0, 8.0, 10.0], [9.0, 5.0, 10.0]]),
                        np.array([[14.0, 8.0, 10.0], [11.0, 12.0, 10.0]]))
    sf = np.array([1.0, 0.8 - 0.2j, 0.5 + 0.1j])
    Q = oopt.source_matrix(om, op["s_id"], op["s_w"], sf)
    QA = oopt.source_matrix(om, op["d_id"], op["d_w"], sf)
    link = np.array([[0, 1], [1, 0], [1, 1], [0, 0]])
    return dict(m=m, om=om, df=df, Q=Q, QA=QA, op=op, phi=osol.direct(A, Q), phiA=osol.direct(A, QA),
                link=link, sf=sf)


def test_jacobian_links_fd_and_reciprocity(two_by_two):
    """Every link (source s_m, detector d_m) of a 2 x 2 set: the exact
    derivative against central differences of meas_flu, and the
    nodal-product J_complex within the approximation's 15 % of it; with
    the two columns of the link table swapped both fail.  Reciprocity:
    the forward field read at a detector equals the adjoint field read at
    the source, d_m . phi_s = q_s . phiA_d."""
    from oracle import fem, jacobian as ojac, optode as oopt, solver as osol
    P = two_by_two
    m, om, link = P["m"], P["om"], P["link"]
    data = oopt.meas_flu(om, P["op"]["d_id"], P["op"]["d_w"], link, P["phi"])
    for q, (s, d) in enumerate(link):
        qs = P["Q"][:, :, s] / P["sf"][None, :]
        recip = (qs * P["phiA"][:, :, d]).sum(axis=0)
        assert np.abs(recip - data[q]).max() < 1e-12 * np.abs(data[q]).max()
    nodes = [_node(m, p) for p in [(10.0, 8.0, 8.0), (7.0, 7.0, 9.0), (12.0, 10.0, 8.0)]]
    Je = ojac.j_exact(om, P["phi"], P["phiA"], link, P["sf"], nodes)
    Jn = ojac.j_complex(P["phi"], P["phiA"], link, P["sf"], 1.0)
    Jswap = ojac.j_complex(P["phi"], P["phiA"], link[:, ::-1], P["sf"], 1.0)

    def meas(mua):
        om2 = fem.OMesh(node=om.node, elem=om.elem, nr=om.nr, mua=mua, musp=om.musp, kappa=om.kappa.copy())
        fem.prepare_mesh(om2)
        om2.K, om2.M = fem.gen_fwdmat(om2)
        ph = osol.direct(fem.full_system(om2, P["df"]), P["Q"])
        return oopt.meas_flu(om2, P["op"]["d_id"], P["op"]["d_w"], link, ph)
    swap_bad = 0
    for q, v in enumerate(nodes):
        dlt = 1e-5
        mp, mm = om.mua.copy(), om.mua.copy()
        mp[v] += dlt
        mm[v] -= dlt
        fd = (meas(mp) - meas(mm)) / (2 * dlt)                       # NM x NF
        scale = np.abs(Je[:, :, q]).max()
        assert np.abs(fd - Je[:, :, q]).max() < 1e-6 * scale
        assert (np.abs(Jn[:, :, v] - Je[:, :, q]) <= 0.15 * np.abs(Je[:, :, q])).all()
        swap_bad += int((np.abs(Jswap[:, :, v] - Je[:, :, q]) > 0.15 * np.abs(Je[:, :, q])).any())
    assert swap_bad == len(nodes)
