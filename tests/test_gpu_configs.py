"""The plain single-frequency forward case at its full size (the reference's
CPU-runnable configuration): 60 x 60 x 30 mm slab on a 2 mm Kuhn grid (15,376
nodes, 81,000 tetrahedra, interior nodes jittered 0.2 mm), mua 0.01/mm,
musp 1/mm, n = 1.4, 16 sources and 16 detectors on 4 x 4 top-face grids, all
256 links, a 10 ns time axis (DC + 100 MHz, the reference solves every
frequency from 0), complex128 solve to 1e-10.  The HIP path (placement,
assembly, FSAI, solver in the reference's BiCGSTAB and in COCG, detector
data) against the FP64 oracle's placement and direct sparse solve."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def config1():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import meshgen, mesh as dm
    from oracle import fem, optode as oopt, solver as osol, transforms as otr
    m, src, det, link = meshgen.slab_60x60x30_h2()
    t = np.arange(0, 501) * 20e-12
    nf = 2
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
    opt = dm.DOptode(src, det, link, 1e3, meshgen.gaussian_tpsf(t, 250e-12), t, nf)
    D.add_optode(opt)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    f, df = otr.freq_axis(t, nf)
    A = fem.full_system(om, df)
    snap = oopt.snap2mesh(om, src, det)
    Q = oopt.source_matrix(om, snap["s_id"], snap["s_w"], opt.s_fpsf)
    phi = osol.direct(A, Q)
    data = oopt.meas_flu(om, snap["d_id"], snap["d_w"], link, phi)
    return dict(D=D, opt=opt, om=om, snap=snap, Q=Q, phi=phi, data=data, link=link, f=f)


def test_config1_placement(config1):
    """Nodal source and detector weights of the GPU placement equal the
    oracle's (element choice on shared faces may differ; the nodal weights
    may not)."""
    from oracle import optode as oopt
    om, opt, snap = config1["om"], config1["opt"], config1["snap"]
    assert config1["D"].nn == 15376 and om.elem.shape[0] == 81000
    for (eid, w), oid, ow in ((opt.s_bary, snap["s_id"], snap["s_w"]), (opt.d_bary, snap["d_id"], snap["d_w"])):
        got = oopt.source_matrix(om, np.asarray(eid), np.asarray(w), np.ones(1))
        ref = oopt.source_matrix(om, oid, ow, np.ones(1))
        assert np.abs(got - ref).max() < 1e-12


@pytest.mark.parametrize("mode,method,bound", [("fp64", "bicgstab", 1e-8), ("plain", "bicgstab", 1e-8),
                                               ("mixed", "cocg", 1e-7)])
def test_config1_forward_vs_direct(config1, mode, method, bound):
    D, phi, data = config1["D"], config1["phi"], config1["data"]
    print(f"[case] config 1 forward, {mode} {method}", flush=True)
    f = D.solve_point_sources("s", mode=mode, tol=1e-10, method=method)
    assert (f.status == 0).all(), f.status
    assert (f.relres < 1e-10).all(), f.relres.max()
    got = f.to_reference_layout().cpu().numpy()                  # NN x NF x NS
    for k in range(got.shape[1]):
        for s in range(got.shape[2]):
            assert rel(got[:, k, s], phi[:, k, s]) < bound, (mode, k, s)
    meas = D.meas_flu(f).cpu().numpy()                           # 256 x NF
    assert meas.shape == (256, 2)
    assert rel(meas, data) < bound
