"""The committed fixture tests/golden/tiny_slab.npz (inputs and FP64 oracle
outputs of one small case of the whole path, written by
tests/golden/make_golden.py).

CPU: the oracle still reproduces the fixture (a pin against later edits of
the oracle), and the fixture satisfies identities the mathematics fixes.
GPU: the HIP path reproduces the fixture, every step through the C ABI."""
import importlib.util
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "tiny_slab.npz")
INPUTS = ["node", "elem", "mua", "musp", "nr", "s_pos", "d_pos", "link", "t", "tpsf", "nf", "avg_pow", "gates",
          "dxyz"]


def load():
    z = np.load(FIXTURE)          # arrays only: allow_pickle stays False
    return {k: z[k] for k in z.files}


def test_oracle_reproduces_fixture():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    g = load()
    y = mg.outputs({k: g[k] for k in INPUTS})
    assert set(y) | set(INPUTS) == set(g)
    for k, v in y.items():
        if np.issubdtype(np.asarray(v).dtype, np.integer):
            assert np.array_equal(v, g[k]), k
        else:
            assert rel(v, g[k]) < 1e-12, k


def test_fixture_identities():
    """1'M1 = box volume; rows of the barycentric weights sum to 1; fc2tg of
    the data equals (Nt/(Nt-1)) sum_t gate(t) fc2td(data)(t) on the fixture's
    own time axis (t = 0:dt:(Nt-1)dt)."""
    from oracle import transforms as otr
    g = load()
    n = g["node"].shape[0]
    M = sp.csr_matrix((g["M_data"], g["M_indices"], g["M_indptr"]), shape=(n, n))
    assert abs(M.sum() - 12.0 * 12.0 * 6.0) < 1e-9
    assert np.allclose(g["s_w"].sum(1), 1.0) and np.allclose(g["d_w"].sum(1), 1.0)
    t = g["t"]
    td_full = otr.fc2td(g["data"], g["f_axis"], t, 1)                 # NM x NT
    tg = t.size / (t.size - 1) * td_full @ g["gates"].T
    assert rel(tg, g["tg"]) < 1e-12


@pytest.mark.gpu
def test_hip_path_reproduces_fixture():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import mesh as dm, transforms as tr
    from dogpup_amd.jacobian import jacobian_tg
    g = load()
    n, nf = g["node"].shape[0], int(g["nf"])
    print("[case] golden tiny slab", flush=True)
    D = dm.DMesh(g["node"], g["elem"], float(g["nr"]), np.stack([g["mua"], g["musp"]], 1))
    rA, cA, K, M = D.system_csr()
    Kg = sp.csr_matrix((K.cpu().numpy(), cA.cpu().numpy(), rA.cpu().numpy()), shape=(n, n))
    Mg = sp.csr_matrix((M.cpu().numpy(), cA.cpu().numpy(), rA.cpu().numpy()), shape=(n, n))
    Kf = sp.csr_matrix((g["K_data"], g["K_indices"], g["K_indptr"]), shape=(n, n))
    Mf = sp.csr_matrix((g["M_data"], g["M_indices"], g["M_indptr"]), shape=(n, n))
    assert abs(Kg - Kf).max() <= 1e-13 * abs(Kf).max()
    assert abs(Mg - Mf).max() <= 1e-13 * abs(Mf).max()
    opt = dm.DOptode(g["s_pos"], g["d_pos"], g["link"], float(g["avg_pow"]), g["tpsf"], g["t"], nf)
    D.add_optode(opt)
    D.dxyz = float(g["dxyz"])
    assert rel(opt.s_fpsf, g["s_fpsf"]) < 1e-13
    for (eid, w), fid, fw in ((opt.s_bary, g["s_id"], g["s_w"]), (opt.d_bary, g["d_id"], g["d_w"])):
        assert np.array_equal(np.asarray(eid), fid)
        assert np.abs(np.asarray(w) - fw).max() < 1e-12
    ns, nd = len(g["s_id"]), len(g["d_id"])
    S = D.solver(ns + nd, "fp64")
    S.ensure_matrix()
    assert np.array_equal(np.asarray(D.g_pattern[0]), g["G_rptr"])
    assert np.array_equal(np.asarray(D.g_pattern[1]), g["G_cptr"])
    assert rel(S.fsai_values().cpu().numpy().T, g["G_val"]) < 1e-12
    for mode, tol, bound in (("fp64", 1e-12, 1e-9), ("mixed", 1e-10, 1e-7)):
        f = D.solve_point_sources("sd", mode=mode, tol=tol, method="cocg")
        assert (f.status == 0).all(), (mode, f.status)
        x = f.to_reference_layout().cpu().numpy()
        assert rel(x[:, :, :ns], g["phi"]) < bound, mode
        assert rel(x[:, :, ns:], g["psi"]) < bound, mode
    fw = D.solve_point_sources("sd", mode="fp64", tol=1e-12, method="cocg")
    fs = D.solve_point_sources("s", mode="fp64", tol=1e-12, method="cocg")
    data = D.meas_flu(fs)
    assert rel(data.cpu().numpy(), g["data"]) < 1e-9
    f_axis, t = g["f_axis"], g["t"]
    dref = torch.as_tensor(g["data"])
    assert rel(tr.fc2td(dref, f_axis, t[:200], 1).cpu().numpy(), g["td"]) < 1e-12
    gfc = tr.td2fc(torch.as_tensor(g["gates"]), f_axis, t, 1)
    assert rel(gfc.cpu().numpy(), g["gate_fc"]) < 1e-13
    assert rel(tr.fc2tg(dref, g["gate_fc"], f_axis, t.size, 1).cpu().numpy(), g["tg"]) < 1e-12
    J, _ = jacobian_tg(D, torch.as_tensor(g["gate_fc"]), mode="fp64", tol=1e-12, field=fw)   # [v][g][m]
    assert rel(np.transpose(J.cpu().numpy(), (2, 1, 0)), g["jtg"]) < 1e-5
