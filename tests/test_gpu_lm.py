"""Weighted LM reconstruction (wghtd_LM_recon) on a small slab: data from an
FP64 forward model with an absorbing inclusion, reconstruction from the
homogeneous background in the mixed and FP64 forms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _problem(mua_field):
    from dogpup_amd import meshgen, mesh as dm
    m = meshgen.slab((20.0, 20.0, 10.0), 1.0, 0.01, 1.0, 1.4)
    t = np.arange(0, 501) * 20e-12
    src = meshgen.top_grid(m.extent, 2, 2, 6.0, 6.0, 8.0, 8.0)
    det = meshgen.top_grid(m.extent, 3, 3, 4.0, 4.0, 6.0, 6.0)
    link = meshgen.all_pairs(4, 9)
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([mua_field(m.node), m.musp], 1))
    D.add_optode(dm.DOptode(src, det, link, 1e3, meshgen.gaussian_tpsf(t, 200e-12), t, 48))
    D.dxyz = 1.0
    gates = np.stack([((t >= a) & (t < a + 500e-12)).astype(float) for a in 200e-12 + 200e-12 * np.arange(8)])
    return D, gates


@pytest.mark.parametrize("kind,w", [("gauss", 1.5), ("square", 4.0)])
def test_target_spots_matches_oracle(kind, w):
    """lm.target_spots (GPU, FP64) against the oracle's restatement of
    dMesh.m:250-282 on a jittered grid with an in-mesh mask: every entry to
    1e-12 of the largest; each row's in-mesh entries sum to minus their
    share of the spot."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import lm
    from oracle import lm as olm
    rng = np.random.default_rng(3)
    g = np.stack(np.meshgrid(np.arange(0.0, 20.0, 1.0), np.arange(0.0, 16.0, 1.0), np.arange(0.0, 8.0, 1.0),
                             indexing="ij"), -1).reshape(-1, 3) + rng.uniform(-0.2, 0.2, (2560, 3))
    inmesh = rng.uniform(size=len(g)) < 0.8
    pos = rng.uniform([2.0, 2.0, 1.0], [18.0, 14.0, 7.0], (7, 3))
    print(f"[case] target_spots {kind}", flush=True)
    got = lm.target_spots(g, pos, w, kind, inmesh).cpu().numpy()
    ref = olm.target_spots(g, pos, w, kind, inmesh)
    assert got.shape == ref.shape == (7, len(g))
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(got[:, ~inmesh] == 0.0)


def test_lm_reconstruction_mixed_and_fp64():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import lm
    from dogpup_amd.jacobian import tg_data
    from dogpup_amd.transforms import td2fc
    centre = np.array([10.0, 10.0, 6.5])
    inc = lambda x: np.where(np.linalg.norm(x - centre, axis=1) <= 3.0, 0.03, 0.01)
    Dt, gates = _problem(inc)
    f = Dt.solve_point_sources("s", mode="fp64", tol=1e-12)
    gfc = td2fc(torch.as_tensor(gates), Dt.optode.fAxis, Dt.optode.tAxis, 1)
    data_m = tg_data(Dt, f, gfc).cpu().numpy()
    assert np.all(data_m > 0)
    pos_s = np.stack(np.meshgrid([6.0, 10.0, 14.0], [6.0, 10.0, 14.0], [5.0, 7.5], indexing="ij"), -1).reshape(-1, 3)
    out = {}
    for mode in ("mixed", "fp64"):
        D, _ = _problem(lambda x: np.full(len(x), 0.01))
        r = lm.wghtd_lm_recon(D, data_m, gates, pos_s, (1e2, 1e-6), mode=mode, tol=1e-10 if mode == "mixed" else 1e-12,
                              max_iter=6)
        assert np.all(np.isfinite(r.mua)) and r.count >= 1, r.exit
        assert min(r.errPerc) < r.errPerc[0]
        inside = np.linalg.norm(D.node - centre, axis=1) <= 3.0
        assert r.mua[inside].mean() > 0.0101                  # absorption recovered at the inclusion
        out[mode] = r
    a, b = out["mixed"].mua, out["fp64"].mua
    assert np.linalg.norm(a - b) / np.linalg.norm(b - 0.01 + 1e-30) < 1e-2
    assert out["mixed"].count == out["fp64"].count


def test_lm_matches_oracle_with_two_step_backs():
    """The LM loop on the GPU path (FP64 solves, J-TG on MFMA, data-space
    tik_inv and reshape_J) against the oracle's restatement of
    wghtd_LM_recon.m:23-168 with its FP64 FEM forward model (direct solves,
    J_complex, fc2tg), on the same time-gated data.  Both forward models are
    wrapped so that the third and fourth evaluations return data scaled by
    3 and 6: the misfit rises twice in a row, so the loop takes the
    step-back branch (:108-116) twice in a row, reusing the cached J .* G'
    and y of the last accepted step.  errPerc, the branch, lambda and mua
    are compared after every update."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import lm, meshgen, mesh as dm
    from oracle import fem, lm as olm, optode as oopt, transforms as otr
    m = meshgen.slab((24.0, 24.0, 12.0), 2.0, 0.01, 1.0, 1.4, jitter=0.2, seed=1,
                     inclusions=[((12.0, 12.0, 7.0), 3.0, 0.1)])
    t = np.arange(0, 1001) * 20e-12
    nf = 64
    src = meshgen.top_grid(m.extent, 3, 3, 5.0, 5.0, 7.0, 7.0)
    det = meshgen.top_grid(m.extent, 3, 3, 8.5, 8.5, 3.5, 3.5)
    link = meshgen.all_pairs(9, 9)
    tpsf = meshgen.gaussian_tpsf(t, 250e-12)
    gates = meshgen.boxcar_gates(t, 8, 500e-12, 200e-12, start=500e-12)
    f, df = otr.freq_axis(t, nf)
    _, s_fpsf = otr.source_spectrum(tpsf, t, f, 1e3)
    gate_fc = otr.td2fc(gates, f, t, 1)
    print("[case] LM parity: oracle data", flush=True)
    om_true = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    snap = oopt.snap2mesh(om_true, src, det)
    data_m, _ = olm.fem_forward(om_true, snap, link, s_fpsf, gate_fc, f, df, t.size, m.h)(m.mua)
    mua0 = np.full(m.node.shape[0], 0.01)
    pos_s = np.array([[12.0, 12.0, 7.0], [6.0, 6.0, 8.0], [18.0, 18.0, 8.0]])
    lam = (10.0, 0.1)
    scale = {2: 3.0, 3: 6.0}                 # forward call -> data factor

    def scripted(fwd):
        calls = [0]

        def wrapped(*a):
            y, J = fwd(*a)
            s = scale.get(calls[0], 1.0)
            calls[0] += 1
            return y * s, J
        return wrapped

    print("[case] LM parity: oracle loop", flush=True)
    om = fem.make_mesh(m.node, m.elem, m.nr, mua0, m.musp)
    fo = olm.fem_forward(om, oopt.snap2mesh(om, src, det), link, s_fpsf, gate_fc, f, df, t.size, m.h)
    ref = olm.wghtd_lm_recon(scripted(fo), mua0, data_m, olm.target_spots(m.node, pos_s, 1.5, "gauss"), lam,
                             max_iter=6)
    print("[case] LM parity: GPU loop", flush=True)
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([mua0, m.musp], 1))
    D.add_optode(dm.DOptode(src, det, link, 1e3, tpsf, t, nf))
    D.dxyz = m.h
    gw = scripted(lambda mesh, run: run(mesh))
    r = lm.wghtd_lm_recon(D, data_m, gates, pos_s, lam, mode="fp64", tol=1e-12, max_iter=6,
                          forward=lambda mesh, run: gw(mesh, run))
    print("oracle:", ref["branch"], [f"{e:.4f}" for e in ref["errPerc"]], ref["exit"], flush=True)
    print("GPU:   ", r.branch, [f"{e:.4f}" for e in r.errPerc], r.exit, flush=True)
    assert ref["branch"][:4] == ["accept", "accept", "step back", "step back"]
    assert r.branch == ref["branch"] and r.exit == ref["exit"] and r.count == ref["count"]
    assert np.allclose(r.errPerc, ref["errPerc"], rtol=1e-5, atol=0)
    assert np.allclose(r.lam, ref["lam"], rtol=1e-5, atol=0)
    for i, (a, b) in enumerate(zip(r.mua_hist, ref["mua_hist"])):
        err = np.linalg.norm(a - b) / np.linalg.norm(b - mua0)
        print(f"update {i + 1}: mua rel diff (to the update size) {err:.2e}", flush=True)
        assert err < 1e-4, (i, err)
    assert np.linalg.norm(r.mua - ref["mua"]) <= 1e-4 * np.linalg.norm(ref["mua"] - mua0)
