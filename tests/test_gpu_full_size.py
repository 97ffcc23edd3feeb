"""Size-independent checks of the batched multi-frequency forward solve at
its full size: the 80 x 80 x 40 mm, 1 mm Kuhn slab (269,001 nodes, 1,536,000
tetrahedra) with a 6 mm sphere of mua 0.03/mm, 32 sources, 64 frequencies
0-3.15 GHz (2048 RHS), mixed COCG with frequency continuation to 5e-8.

The relative residual ||b - A_k x|| / ||b|| of every column at sampled
frequencies (0, 1.6 and 3.15 GHz, solved from x0 = 0, and 50, 100, 200,
400 MHz and 3.05 GHz, started from Galerkin guesses) is computed off the
solver's path: A_k from the oracle's own FP64
assembly (dMesh.m:1033-1244), b from the oracle's source matrix on the
product's placement (dMesh.m:124-129), the product in scipy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 5e-8


def test_full_size_residuals_off_solver_path():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import meshgen, mesh as dm
    from oracle import fem, optode as oopt
    m, src, det, link = meshgen.slab_80x80x40_h1(seed=0)
    t = np.arange(0, 1001) * 20e-12
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
    opt = dm.DOptode(src, det, link, 1e3, meshgen.gaussian_tpsf(t, 250e-12), t, 64)
    D.add_optode(opt)
    assert D.nn == 269001
    print("[case] full size: mixed COCG with continuation, 2048 RHS", flush=True)
    fw = D.solve_point_sources("s", mode="mixed", tol=TOL)
    assert (fw.status == 0).all() and (fw.relres < TOL).all()
    print("[case] full size: oracle assembly and residuals", flush=True)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    A = fem.full_system(om, opt.df)
    eid, w = opt.s_bary
    # level-0 frequencies (solved from x0 = 0) and frequencies of every later
    # continuation level (started from Galerkin guesses): 8, 4, 2, 1 and 61
    ks = (0, 1, 2, 4, 8, 32, 61, 63)
    Q = oopt.source_matrix(om, np.asarray(eid), np.asarray(w), opt.s_fpsf[list(ks)])  # NN x 3 x NS
    X = fw.to_reference_layout()
    for q, k in enumerate(ks):
        Ak = fem.system_at(A, k)
        xk = X[:, k, :].cpu().numpy()
        r = Q[:, q, :] - Ak @ xk
        rel = np.linalg.norm(r, axis=0) / np.linalg.norm(Q[:, q, :], axis=0)
        print(f"frequency {k} ({k * 50} MHz): relres max {rel.max():.3e} (solver's own {fw.relres[k].max():.3e})",
              flush=True)
        assert rel.max() < 1.05 * TOL
        assert np.allclose(rel, fw.relres[k], rtol=1e-3, atol=1e-12)
