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
