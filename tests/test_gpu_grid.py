"""GPU point location, voxel basis (mesh2grid), mesh2mesh, display slices and
the voxel-basis Jacobian and LM step, against the oracle's brute-force
restatement (dMesh.m:286-411, :206-247)."""
import numpy as np
import pytest

from conftest import rel, small_slab

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _lin(p):
    return 0.3 + 1.7 * p[:, 0] - 0.4 * p[:, 1] + 2.2 * p[:, 2]


def test_locator_matches_brute_force():
    _need_gpu()
    from dogpup_amd import grid
    from oracle import grid as og
    m = small_slab(seed=7, extent=(12.0, 10.0, 8.0), h=1.0)
    rng = np.random.default_rng(0)
    pts = np.concatenate([rng.uniform([-1, -1, -1], [13, 11, 9], size=(3000, 3)), m.node,
                          m.node[m.elem].mean(axis=1)[:500]])
    print(f"[case] locator {m.node.shape[0]} nodes, {m.elem.shape[0]} tets, {pts.shape[0]} points", flush=True)
    L = grid.Locator(m.node, m.elem)
    assert L.stats()["pairs"] >= m.elem.shape[0]
    eid, w, nodes = (t.cpu().numpy() for t in L.find(pts))
    eo, wo = og._locate(m.node, m.elem, pts)
    assert np.array_equal(eid >= 0, eo >= 0)
    found = eid >= 0
    assert np.array_equal(nodes[found], m.elem[eid[found]])
    assert (nodes[~found] == -1).all() and np.isnan(w[~found]).all()
    # unambiguous points: same element and weights
    clear = found & (np.nan_to_num(wo, nan=-1).min(axis=1) > 1e-7)
    assert clear.sum() > 1500
    assert np.array_equal(eid[clear], eo[clear])
    assert np.abs(w[clear] - wo[clear]).max() < 1e-12
    # every located point interpolates a nodal field the same way
    f = rng.normal(size=m.node.shape[0])
    gi = (w[found] * f[nodes[found]]).sum(axis=1)
    oi = (wo[found] * f[m.elem[eo[found]]]).sum(axis=1)
    assert np.abs(gi - oi).max() < 1e-12


@pytest.mark.parametrize("case", ["slab", "cylinder"])
def test_mesh2grid_matches_oracle(case):
    _need_gpu()
    from dogpup_amd import grid, meshgen
    from oracle import grid as og
    if case == "slab":
        m = small_slab(seed=4, extent=(12.0, 10.0, 8.0), h=2.0)
        x, y, z = -1.3 + 1.7 * np.arange(9), -0.9 + 1.3 * np.arange(10), -0.7 + 1.1 * np.arange(9)
    else:   # coarse grid on a cylinder: nodes near the curved side take the nearest voxel
        m = meshgen.cylinder(5.0, 6.0, 1.0)
        x = y = np.arange(-0.4, 13.1, 2.5)
        z = np.arange(-0.3, 8.8, 2.5)
    print(f"[case] mesh2grid {case}", flush=True)
    b = grid.mesh2grid(m.node, m.elem, x, y, z)
    r = og.mesh2grid(m.node, m.elem, x, y, z)
    assert np.array_equal(b.gridinmesh, r["gridinmesh"])
    assert b.n_fallback == len(r["fallback"])
    if case == "cylinder":
        assert b.n_fallback > 0
    assert abs(b.m2g.to_scipy() - r["m2g"]).max() < 1e-12
    assert abs(b.g2m.to_scipy() - r["g2m"]).max() < 1e-12
    assert b.dxyz == r["dxyz"] and b.gridSize == r["gridSize"]
    # device application equals the sparse products
    f = np.random.default_rng(1).normal(size=m.node.shape[0])
    assert np.abs(b.to_grid(f).cpu().numpy() - r["m2g"] @ f).max() < 1e-12
    gv = np.random.default_rng(2).normal(size=b.nv)
    assert np.abs(b.to_mesh(gv).cpu().numpy() - r["g2m"] @ gv).max() < 1e-12


def test_grid_to_mesh_random_mask_matches_oracle():
    """In-mesh filter, renormalisation and nearest-voxel fallback under an
    arbitrary in-mesh mask."""
    _need_gpu()
    from dogpup_amd import _lib
    from oracle import grid as og
    m = small_slab(seed=8, extent=(12.0, 10.0, 8.0), h=2.0)
    x, y, z = -1.3 + 1.7 * np.arange(9), -0.9 + 1.3 * np.arange(10), -0.7 + 1.1 * np.arange(9)
    nv = x.size * y.size * z.size
    inm = np.random.default_rng(3).uniform(size=nv) < 0.25
    ref, fb = og.grid_to_mesh(m.node, x, y, z, inm)
    nn = m.node.shape[0]
    idx = torch.empty((nn, 4), dtype=torch.int32, device="cuda")
    w = torch.empty((nn, 4), dtype=torch.float64, device="cuda")
    counts = np.zeros(2, np.int32)
    print(f"[case] grid_to_mesh random mask, {inm.sum()} of {nv} voxels", flush=True)
    _lib.call("dogpup_grid_to_mesh", _lib.ptr(torch.as_tensor(m.node, device="cuda")), nn,
              _lib.ptr(np.array([x[0], y[0], z[0]])), _lib.ptr(np.array([1.7, 1.3, 1.1])),
              _lib.ptr(np.array([x.size, y.size, z.size], np.int32)),
              _lib.ptr(torch.as_tensor(inm.astype(np.uint8), device="cuda")), _lib.ptr(idx), _lib.ptr(w),
              _lib.ptr(counts), _lib.stream_ptr())
    from dogpup_amd.grid import Ell4
    got = Ell4(idx, w, nn, nv).to_scipy()
    assert counts[0] == len(fb) and len(fb) > 0 and counts[1] == 0
    assert abs(got - ref).max() < 1e-12


def test_grid_to_mesh_rejects_nodes_outside_grid():
    _need_gpu()
    from dogpup_amd import grid
    m = small_slab(seed=4, extent=(12.0, 10.0, 8.0), h=2.0)
    with pytest.raises(Exception, match="outside the voxel grid"):
        grid.mesh2grid(m.node, m.elem, np.arange(0, 10.0, 2.0), np.arange(0, 11.0, 2.0), np.arange(0, 9.0, 2.0))


def test_field_to_voxels_matches_sparse_product():
    _need_gpu()
    from dogpup_amd import grid
    from dogpup_amd.mesh import FluenceField
    m = small_slab(seed=4, extent=(12.0, 10.0, 8.0), h=2.0)
    b = grid.mesh2grid(m.node, m.elem, *(np.arange(-0.5, e + 2.6, 2.5) for e in (12.0, 10.0, 8.0)))
    rng = np.random.default_rng(4)
    nn, ng, nsp = m.node.shape[0], 6, 8
    data = torch.as_tensor(rng.normal(size=(ng, nn, nsp)) + 1j * rng.normal(size=(ng, nn, nsp)), device="cuda")
    f = FluenceField(data, nn, 3, 12, nsp)
    v = b.field_to_voxels(f)
    M = b.m2g.to_scipy()[b.in_idx]
    ref = np.stack([M @ data[g].cpu().numpy() for g in range(ng)])
    assert v.data.shape == (ng, b.nv_in, nsp)
    assert np.abs(v.data.cpu().numpy() - ref).max() < 1e-12


def test_mesh2mesh_and_slices():
    _need_gpu()
    from dogpup_amd import grid
    from oracle import grid as og
    old = small_slab(seed=5, extent=(8.0, 8.0, 4.0), h=2.0)
    new = small_slab(seed=6, extent=(8.0, 8.0, 4.0), h=1.0)
    pts = np.concatenate([new.node, [[9.0, 1.0, 1.0], [-0.5, 4.0, 2.0]]])
    val = np.stack([_lin(old.node), np.random.default_rng(5).normal(size=old.node.shape[0])], 1)
    got = grid.mesh2mesh(pts, val, old.node, old.elem).cpu().numpy()
    ref = og.mesh2mesh(pts, val, old.node, old.elem)
    assert np.abs(got - ref).max() < 1e-12
    assert np.allclose(got[: new.node.shape[0], 0], _lin(new.node), atol=1e-12)
    M, p, outside = grid.mesh_slice(old.node, old.elem, "z=2")
    assert p.shape == (256 * 256, 3) and np.allclose(p[:, 2], 2.0)
    vals = M.apply(_lin(old.node)).cpu().numpy()
    ins = np.all((p >= 0) & (p <= [8.0, 8.0, 4.0]), axis=1)
    assert np.array_equal(~outside, ins)
    assert np.allclose(vals[ins], _lin(p[ins]), atol=1e-12)
    M3, p3 = grid.mesh_slice3(old.node, old.elem)
    assert p3.shape == (64 ** 3, 3)
    v3 = M3.apply(_lin(old.node)).cpu().numpy()
    ins3 = np.all((p3 >= 0) & (p3 <= [8.0, 8.0, 4.0]), axis=1)
    assert np.allclose(v3[ins3], _lin(p3[ins3]), atol=1e-12) and not v3[~ins3].any()


def test_jacobian_tg_on_voxel_basis():
    """J on a 2.5 mm voxel grid: fields interpolated by m2g, then the TG
    contraction, against the oracle's J_complex on m2g-interpolated direct
    solves followed by fc2tg."""
    _need_gpu()
    from dogpup_amd import meshgen, mesh as dm
    from dogpup_amd.jacobian import jacobian_tg
    from oracle import fem, grid as og, jacobian as ojac, optode as oopt, solver as osol, transforms as otr
    m = small_slab(seed=3, inclusions=[((10.0, 10.0, 5.0), 3.0, 0.03)])
    t = np.arange(0, 501) * 20e-12
    nf = 4
    src = meshgen.top_grid(m.extent, 2, 2, 5.0, 5.0, 10.0, 10.0)
    det = meshgen.top_grid(m.extent, 3, 2, 4.0, 7.0, 6.0, 6.0)
    link = meshgen.all_pairs(4, 6)
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
    opt = dm.DOptode(src, det, link, 1e3, meshgen.gaussian_tpsf(t, 250e-12), t, nf)
    D.add_optode(opt)
    x = y = np.arange(-0.25, 23.0, 2.5)
    z = np.arange(-0.25, 12.3, 2.5)
    D.mesh2grid(x, y, z)
    print(f"[case] voxel-basis J, {D.basis.nv_in} in-mesh voxels of {D.basis.nv}", flush=True)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    f, df = otr.freq_axis(t, nf)
    A = fem.full_system(om, df)
    sid, sw = opt.s_bary
    did, dw = opt.d_bary
    phi = osol.direct(A, oopt.source_matrix(om, sid, sw, opt.s_fpsf))
    phiA = osol.direct(A, oopt.source_matrix(om, did, dw, opt.s_fpsf))
    r = og.mesh2grid(m.node, m.elem, x, y, z)
    nn = m.node.shape[0]
    pg = (r["m2g"] @ phi.reshape(nn, -1)).reshape(-1, nf, phi.shape[2])
    pAg = (r["m2g"] @ phiA.reshape(nn, -1)).reshape(-1, nf, phiA.shape[2])
    gates = np.stack([((t >= a) & (t < a + 500e-12)).astype(float) for a in (0.0, 200e-12, 400e-12)])
    gfc = otr.td2fc(gates, f, t, 1)
    ref = ojac.j_tg(pg, pAg, link, opt.s_fpsf, r["dxyz"], gfc, f, t.size)[:, :, r["gridinmesh"]]
    J, _ = jacobian_tg(D, torch.as_tensor(gfc), mode="fp64", tol=1e-12)
    assert J.shape == (D.basis.nv_in, gfc.shape[0], link.shape[0])
    assert rel(np.transpose(J.cpu().numpy(), (2, 1, 0)), ref) < 1e-5
    # back on the nodal basis, the mesh's own dxyz (unset here) is restored
    assert D.dxyz == r["dxyz"]
    D.clear_grid()
    assert D.basis is None and D.dxyz is None


def test_lm_on_voxel_basis():
    """A few weighted LM iterations with the unknowns on 2 mm voxels over a
    1 mm mesh: finite, error decreasing, absorption raised at the inclusion."""
    _need_gpu()
    from dogpup_amd import lm, meshgen, mesh as dm
    from dogpup_amd.jacobian import tg_data
    from dogpup_amd.transforms import td2fc
    centre = np.array([10.0, 10.0, 6.5])

    def problem(mua_field):
        m = meshgen.slab((20.0, 20.0, 10.0), 1.0, 0.01, 1.0, 1.4)
        t = np.arange(0, 501) * 20e-12
        src = meshgen.top_grid(m.extent, 2, 2, 6.0, 6.0, 8.0, 8.0)
        det = meshgen.top_grid(m.extent, 3, 3, 4.0, 4.0, 6.0, 6.0)
        D = dm.DMesh(m.node, m.elem, m.nr, np.stack([mua_field(m.node), m.musp], 1))
        D.add_optode(dm.DOptode(src, det, meshgen.all_pairs(4, 9), 1e3, meshgen.gaussian_tpsf(t, 200e-12), t, 48))
        gates = np.stack([((t >= a) & (t < a + 500e-12)).astype(float) for a in 200e-12 + 200e-12 * np.arange(8)])
        return D, gates

    Dt, gates = problem(lambda x: np.where(np.linalg.norm(x - centre, axis=1) <= 3.0, 0.03, 0.01))
    f = Dt.solve_point_sources("s", mode="fp64", tol=1e-12)
    gfc = td2fc(torch.as_tensor(gates), Dt.optode.fAxis, Dt.optode.tAxis, 1)
    data_m = tg_data(Dt, f, gfc).cpu().numpy()
    D, _ = problem(lambda x: np.full(len(x), 0.01))
    D.mesh2grid(np.arange(0, 21.0, 2.0), np.arange(0, 21.0, 2.0), np.arange(0, 11.0, 2.0))
    assert D.basis.gridinmesh.all() and D.basis.n_fallback == 0
    pos_s = np.stack(np.meshgrid([6.0, 10.0, 14.0], [6.0, 10.0, 14.0], [5.0, 7.5], indexing="ij"), -1).reshape(-1, 3)
    print("[case] LM on a 2 mm voxel basis", flush=True)
    r = lm.wghtd_lm_recon(D, data_m, gates, pos_s, (1e2, 1e-1), mode="mixed", tol=1e-10, max_iter=4,
                          method="cocg")
    assert np.all(np.isfinite(r.mua)) and r.count >= 1, r.exit
    assert min(r.errPerc) < r.errPerc[0]
    inside = np.linalg.norm(D.node - centre, axis=1) <= 3.0
    assert r.mua[inside].mean() > 0.0101
