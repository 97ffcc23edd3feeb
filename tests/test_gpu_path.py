"""GPU parity of the HIP path against the FP64 oracle (small meshes).

Every call goes through libdogpup_hip.so's C ABI; every result is compared
with the oracle's.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel, small_slab

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def setup():
    _need_gpu()
    from dogpup_amd import meshgen, mesh as dm
    from oracle import fem, optode as oopt, transforms as otr
    m = small_slab(seed=3, inclusions=[((10.0, 10.0, 5.0), 3.0, 0.03)])
    t = np.arange(0, 501) * 20e-12
    nf = 4
    src = meshgen.top_grid(m.extent, 2, 2, 5.0, 5.0, 10.0, 10.0)
    det = meshgen.top_grid(m.extent, 3, 2, 4.0, 7.0, 6.0, 6.0)
    link = meshgen.all_pairs(4, 6)
    tpsf = meshgen.gaussian_tpsf(t, 250e-12)
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
    opt = dm.DOptode(src, det, link, 1e3, tpsf, t, nf)
    D.add_optode(opt)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    f, df = otr.freq_axis(t, nf)
    A = fem.full_system(om, df)
    return dict(m=m, D=D, opt=opt, om=om, A=A, t=t, f=f, df=df, nf=nf, link=link, src=src, det=det)


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to("cuda", dtype)


def test_assembly_matches_oracle(setup):
    D, A = setup["D"], setup["A"]
    om = setup["om"]
    rA, cA, K, M = D.system_csr()
    Kg = sp.csr_matrix((K.cpu().numpy(), cA.cpu().numpy(), rA.cpu().numpy()), shape=A.shape)
    Mg = sp.csr_matrix((M.cpu().numpy(), cA.cpu().numpy(), rA.cpu().numpy()), shape=A.shape)
    Kf = (om.K + sp.triu(om.K, 1).T).tocsr()
    Mf = (om.M + sp.triu(om.M, 1).T).tocsr()
    assert abs(Kg - Kf).max() <= 1e-13 * abs(Kf).max()
    assert abs(Mg - Mf).max() <= 1e-13 * abs(Mf).max()


def test_compat_smv_fsaimv(setup):
    from dogpup_amd import _lib
    from oracle import fsai as ofsai, solver as osol
    A = setup["A"].tocsr()
    A.sort_indices()
    n, nf, ns = A.shape[0], 3, 5
    rng = np.random.default_rng(0)
    X = rng.normal(size=(n, ns, nf)) + 1j * rng.normal(size=(n, ns, nf))
    f = np.array([0.0, 1.0, 2.5])
    ref = osol.smv(A, f, X)
    # planar device launcher (smv_call), reference layout x(row, col, freq)
    xr = dev(np.transpose(X, (2, 1, 0)).real.copy())
    xi = dev(np.transpose(X, (2, 1, 0)).imag.copy())
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    _lib.call("dogpup_smv_call", _lib.ptr(yr), _lib.ptr(yi), _lib.ptr(xr), _lib.ptr(xi), _lib.ptr(dev(f)),
              _lib.ptr(dev(A.data.real)), _lib.ptr(dev(A.data.imag)), _lib.ptr(dev(A.indptr, torch.int32)),
              _lib.ptr(dev(A.indices, torch.int32)), n, ns, nf, None)
    torch.cuda.synchronize()
    got = np.transpose(yr.cpu().numpy() + 1j * yi.cpu().numpy(), (2, 1, 0))
    assert rel(got, ref) < 1e-14
    # mex-level (host interleaved, MATLAB column-major NN x NS x NF)
    xh = np.asfortranarray(X).ravel(order="F")
    yh = np.zeros(2 * xh.size)
    Ah = np.ascontiguousarray(A.data)
    _lib.call("dogpup_smv_mex", _lib.ptr(yh), _lib.ptr(xh.view(np.float64)), _lib.ptr(f), _lib.ptr(Ah.view(np.float64)),
              _lib.ptr(A.indptr.astype(np.int32)), _lib.ptr(A.indices.astype(np.int32)), n, ns, nf, A.nnz)
    got2 = yh.view(np.complex128).reshape((n, ns, nf), order="F")
    assert rel(got2, ref) < 1e-14
    # FSAImv with per-frequency values
    rG, cG, valG = ofsai.fsaip_gen_batched(A, nf)
    refG = osol.fsaimv(rG, cG, valG, X)
    Gh = np.asfortranarray(valG).ravel(order="F")
    yh = np.zeros(2 * xh.size)
    _lib.call("dogpup_fsaimv_mex", _lib.ptr(yh), _lib.ptr(xh.view(np.float64)), _lib.ptr(Gh.view(np.float64)),
              _lib.ptr(rG), _lib.ptr(cG), n, ns, nf, cG.size)
    got3 = yh.view(np.complex128).reshape((n, ns, nf), order="F")
    assert rel(got3, refG) < 1e-14


def test_compat_device_launchers_fsaimv_fsaip(setup):
    """The device launchers behind FSAImv_mex and FSAIP_gen_mex
    (FSAImv_call, FSAImvMex.hpp:4-15 / FSAImv_CUDA.cu:8-53, and FSAIP_call,
    FSAIP.hpp:4-13 / FSAIP_CUDA.cu:162-180), called directly on planar device
    arrays in the reference layout, against the oracle: y(:,c,k) = G_k x(:,c,k)
    with G_k at valG[nnz*k + i], and the factors of every frequency at
    out[k + nfreqs*j]."""
    from dogpup_amd import _lib
    from oracle import fsai as ofsai, solver as osol
    A = setup["A"].tocsr()
    A.sort_indices()
    n, nf, ns = A.shape[0], 3, 4
    rng = np.random.default_rng(1)
    X = rng.normal(size=(n, ns, nf)) + 1j * rng.normal(size=(n, ns, nf))
    rG, cG, valG = ofsai.fsaip_gen(A, nf)                # per-row partial-pivot GE (the reference's)
    nnz = cG.size
    ref = osol.fsaimv(rG, cG, valG, X)
    xr = dev(np.transpose(X, (2, 1, 0)).real.copy())
    xi = dev(np.transpose(X, (2, 1, 0)).imag.copy())
    gv = np.ascontiguousarray(valG.T)                     # [k][i]: nnz*k + i
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    _lib.call("dogpup_fsaimv_call", _lib.ptr(yr), _lib.ptr(yi), _lib.ptr(xr), _lib.ptr(xi), _lib.ptr(dev(gv.real)),
              _lib.ptr(dev(gv.imag)), _lib.ptr(dev(rG, torch.int32)), _lib.ptr(dev(cG, torch.int32)), nnz, n, ns, nf,
              None)
    torch.cuda.synchronize()
    got = np.transpose(yr.cpu().numpy() + 1j * yi.cpu().numpy(), (2, 1, 0))
    assert rel(got, ref) < 1e-14
    # FSAIP_call: tril(A) in CSR, factors of k = 0..nf-1
    At = sp.tril(A).tocsr()
    At.sort_indices()
    outr = torch.full((nf * nnz,), float("nan"), dtype=torch.float64, device="cuda")
    outi = torch.full_like(outr, float("nan"))
    _lib.call("dogpup_fsaip_call", _lib.ptr(outr), _lib.ptr(outi), _lib.ptr(dev(At.data.real)),
              _lib.ptr(dev(At.data.imag)), _lib.ptr(dev(At.indptr, torch.int32)), _lib.ptr(dev(At.indices, torch.int32)),
              _lib.ptr(dev(rG, torch.int32)), _lib.ptr(dev(cG, torch.int32)), nf, n, None)
    torch.cuda.synchronize()
    got = (outr.cpu().numpy() + 1j * outi.cpu().numpy()).reshape(nnz, nf)        # out[k + nf*j]
    assert rel(got, valG) < 1e-13


def test_compat_fsaip_gen(setup):
    """FSAIP_gen_mex vs the oracle's per-row partial-pivot GE (brute force)."""
    from dogpup_amd import _lib
    from oracle import fsai as ofsai
    A = setup["A"]
    nf = 3
    rG, cG, valG = ofsai.fsaip_gen(A, nf)
    At = sp.tril(A).tocsr()
    At.sort_indices()
    out = np.zeros(2 * nf * cG.size)
    vA = np.ascontiguousarray(At.data)
    _lib.call("dogpup_fsaip_gen_mex", _lib.ptr(out), _lib.ptr(vA.view(np.float64)),
              _lib.ptr(At.indptr.astype(np.int32)), _lib.ptr(At.indices.astype(np.int32)), _lib.ptr(rG),
              _lib.ptr(cG), A.shape[0], At.nnz, cG.size, nf)
    got = out.view(np.complex128).reshape((nf, cG.size), order="F").T
    assert rel(got, valG) < 1e-13


@pytest.mark.parametrize("mode,tol,bound,method", [("fp64", 1e-12, 1e-9, "bicgstab"),
                                                   ("mixed", 1e-10, 1e-7, "bicgstab"),
                                                   ("plain", 1e-12, 1e-9, "bicgstab"),
                                                   ("fp64", 1e-12, 1e-9, "cocg"),
                                                   ("mixed", 1e-10, 1e-7, "cocg")])
def test_solver_vs_direct(setup, mode, tol, bound, method):
    from oracle import optode as oopt, solver as osol
    D, A, om = setup["D"], setup["A"], setup["om"]
    opt = setup["opt"]
    sid, sw = opt.s_bary
    Q = oopt.source_matrix(om, sid, sw, opt.s_fpsf)
    ref = osol.direct(A, Q)
    print(f"[case] solver {mode} {method} tol {tol}", flush=True)
    phi = D.solve_point_sources("s", mode=mode, tol=tol, method=method)
    got = phi.to_reference_layout().cpu().numpy()
    assert (phi.status == 0).all(), phi.status
    assert (phi.relres < tol).all(), phi.relres
    for k in range(setup["nf"]):
        for s in range(Q.shape[2]):
            assert rel(got[:, k, s], ref[:, k, s]) < bound, (mode, method, k, s)


def test_solver_fsai_factors(setup):
    from oracle import fsai as ofsai
    D, A = setup["D"], setup["A"]
    S = D.solver(4, "fp64")
    S.ensure_matrix()
    nf = setup["nf"]
    fscale = np.arange(nf) * setup["df"] / D.c
    rG, cG, valG = ofsai.fsaip_gen_batched(sp.csr_matrix(A.real) + 1j * sp.csr_matrix(A.imag / (setup["df"] / D.c)), nf, fscale=fscale)
    got = S.fsai_values().cpu().numpy().T
    assert (np.asarray(D.g_pattern[0]) == rG).all() and (np.asarray(D.g_pattern[1]) == cG).all()
    assert rel(got, valG) < 1e-12


@pytest.mark.parametrize("mode,tol,bound,method", [("mixed", 1e-10, 1e-7, "cocg"), ("fp64", 1e-12, 1e-9, "cocg"),
                                                   ("plain", 1e-12, 1e-9, "bicgstab")])
def test_solver_on_morton_sorted_mesh_vs_direct(setup, mode, tol, bound, method):
    """A mesh whose nodes already follow the solver's Morton order
    (meshgen.sorted_mesh): the handle works in the caller's b and x with no
    row permutation, and every form still solves to the oracle's direct
    solution of that mesh."""
    from dogpup_amd import meshgen, mesh as dm
    from oracle import fem, optode as oopt, solver as osol
    ms = meshgen.sorted_mesh(setup["m"])
    D = dm.DMesh(ms.node, ms.elem, ms.nr, np.stack([ms.mua, ms.musp], 1))
    assert np.array_equal(D.solver_order, np.arange(D.nn))
    D.add_optode(dm.DOptode(setup["src"], setup["det"], setup["link"], 1e3,
                            meshgen.gaussian_tpsf(setup["t"], 250e-12), setup["t"], setup["nf"]))
    om = fem.make_mesh(ms.node, ms.elem, ms.nr, ms.mua, ms.musp)
    A = fem.full_system(om, setup["df"])
    sid, sw = D.optode.s_bary
    ref = osol.direct(A, oopt.source_matrix(om, sid, sw, D.optode.s_fpsf))
    print(f"[case] sorted mesh {mode} {method}", flush=True)
    phi = D.solve_point_sources("s", mode=mode, tol=tol, method=method)
    got = phi.to_reference_layout().cpu().numpy()
    assert (phi.status == 0).all(), phi.status
    for k in range(setup["nf"]):
        for c in range(ref.shape[2]):
            assert rel(got[:, k, c], ref[:, k, c]) < bound, (mode, k, c)


@pytest.mark.parametrize("mode,method", [("mixed", "cocg"), ("fp64", "cocg"), ("fp64", "bicgstab"),
                                         ("plain", "bicgstab")])
def test_sorted_mesh_new_x_buffer_and_in_place(setup, mode, method):
    """On a mesh in the solver's order the solves iterate in the caller's x,
    so the FP64 fused form's captured iteration holds that address: a second
    solve into another buffer must recapture, not replay into the first one.
    A solve with x = b (in place) reads b before writing x.  Each of the three
    solves gives the first one's field bit for bit."""
    from dogpup_amd import meshgen, mesh as dm
    ms = meshgen.sorted_mesh(setup["m"])
    D = dm.DMesh(ms.node, ms.elem, ms.nr, np.stack([ms.mua, ms.musp], 1))
    D.add_optode(dm.DOptode(setup["src"], setup["det"], setup["link"], 1e3,
                            meshgen.gaussian_tpsf(setup["t"], 250e-12), setup["t"], setup["nf"]))
    S = D.solver(4, mode, method)
    rng = np.random.default_rng(11)
    shp = S.empty_block().shape
    b = torch.as_tensor(rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).to("cuda")
    tol = 1e-10 if mode == "mixed" else 1e-12
    print(f"[case] sorted mesh, new x and in place: {mode} {method}", flush=True)
    f1 = S.solve(b, tol=tol)
    x1 = f1.data.clone()
    f2 = S.solve(b, tol=tol, x=torch.full_like(b, 7.0))   # another buffer, held while x1 is alive
    bi = b.clone()
    f3 = S.solve(bi, tol=tol, x=bi)
    torch.cuda.synchronize()
    assert (f1.status == 0).all(), f1.status
    real = lambda t: torch.view_as_real(t[..., :4]).contiguous().view(torch.int64)
    assert torch.equal(real(f2.data), real(x1)), "second buffer"
    assert torch.equal(real(f3.data), real(x1)), "in place"
    assert np.array_equal(f2.iters, f1.iters) and np.array_equal(f3.iters, f1.iters)


@pytest.mark.parametrize("mode", ["mixed", "fp64", "plain"])
def test_solver_fsai_direct_matches_gather(setup, mode, monkeypatch):
    """The solver's FSAI setup through the A-position table, written straight
    into the solver's G and G^T layouts, against the reference-order factor
    gathered into them (DOGPUP_FSAI_GATHER=1): the same factors, so the
    solves agree bit for bit; the reference-order factors the handle returns
    on request are the oracle's."""
    from oracle import fsai as ofsai
    D, A = setup["D"], setup["A"]
    nf = setup["nf"]
    S = D.solver(4, mode)
    rng = np.random.default_rng(7)
    b = torch.as_tensor(rng.standard_normal(S.empty_block().shape) + 1j * rng.standard_normal(S.empty_block().shape)).to("cuda")
    print(f"[case] fsai direct vs gather {mode}", flush=True)
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("DOGPUP_FSAI_GATHER", flag)
        S._version = None
        S.ensure_matrix()
        out[flag] = S.solve(b, tol=1e-10)
        assert (out[flag].status == 0).all()
    bits = [torch.view_as_real(out[f].data[:, :, :4].contiguous()).view(torch.int64) for f in ("0", "1")]
    assert torch.equal(bits[0], bits[1])
    assert (out["0"].iters == out["1"].iters).all()
    fscale = np.arange(nf) * setup["df"] / D.c
    rG, cG, valG = ofsai.fsaip_gen_batched(sp.csr_matrix(A.real) + 1j * sp.csr_matrix(A.imag / (setup["df"] / D.c)), nf, fscale=fscale)
    assert rel(S.fsai_values().cpu().numpy().T, valG) < 1e-12


@pytest.mark.parametrize("mode", ["mixed", "fp64", "plain"])
def test_graphs_kept_across_new_properties(setup, mode):
    """A handle keeps its captured graphs when dogpup_solver_set_matrix brings
    new optical properties at the same frequency scales: its next solve is
    the solve of a fresh handle with those properties, bit for bit; a call
    with other frequency scales clears them, and the solve is again a fresh
    handle's."""
    from dogpup_amd import _lib, meshgen, mesh as dm
    m = setup["m"]
    tpsf = meshgen.gaussian_tpsf(setup["t"], 250e-12)

    def make(mua):
        D = dm.DMesh(m.node, m.elem, m.nr, np.stack([mua, m.musp], 1))
        D.add_optode(dm.DOptode(setup["src"], setup["det"], setup["link"], 1e3, tpsf, setup["t"], setup["nf"]))
        return D
    mua1 = m.mua * 1.3
    print(f"[case] graphs kept, {mode}", flush=True)
    D = make(m.mua.copy())
    D.solve_point_sources("s", mode=mode, tol=1e-10)                     # captures the graphs
    D.update_properties(np.stack([mua1, m.musp], 1))
    again = D.solve_point_sources("s", mode=mode, tol=1e-10)              # same fscale: graphs kept
    fresh = make(mua1).solve_point_sources("s", mode=mode, tol=1e-10)
    assert (again.status == 0).all()
    assert torch.equal(again.data[:, :, :4], fresh.data[:, :, :4])
    assert (again.iters == fresh.iters).all()
    # other frequency scales through the C ABI: graphs cleared
    def with_scale(Dm, f):
        S = Dm.solver(4, mode)
        S.ensure_matrix()
        if mode == "mixed":
            S.set_continuation(True, Dm.optode.s_fpsf)          # as solve_point_sources sets it
        _lib.call("dogpup_solver_set_matrix", S.h, _lib.ptr(Dm.K), _lib.ptr(Dm.M),
                  _lib.ptr(torch.as_tensor(f, dtype=torch.float64, device="cuda")), _lib.stream_ptr())
        nodes, w = Dm._bary_dev("s")
        spec = torch.as_tensor(Dm.optode.s_fpsf, dtype=torch.complex128, device="cuda")
        b = S.empty_block()
        _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), Dm.nn, S.nf, 4, S.nsp,
                  _lib.ptr(nodes), _lib.ptr(w), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
        return S.solve(b, tol=1e-10)
    f2 = np.arange(setup["nf"]) * setup["df"] / D.c * 1.25
    a2 = with_scale(D, f2)
    b2 = with_scale(make(mua1), f2)
    assert (a2.status == 0).all()
    assert torch.equal(a2.data[:, :, :4], b2.data[:, :, :4])
    assert not torch.equal(a2.data[:, :, :4], again.data[:, :, :4])


def test_meas_and_transforms(setup):
    from dogpup_amd import transforms as tr
    from oracle import optode as oopt, solver as osol, transforms as otr
    D, A, om, opt = setup["D"], setup["A"], setup["om"], setup["opt"]
    sid, sw = opt.s_bary
    did, dw = opt.d_bary
    Q = oopt.source_matrix(om, sid, sw, opt.s_fpsf)
    ref_phi = osol.direct(A, Q)
    ref = oopt.meas_flu(om, did, dw, setup["link"], ref_phi)
    phi, data = D.flu_solve(mode="fp64")
    assert rel(data.cpu().numpy(), ref) < 1e-9
    t, f = setup["t"], setup["f"]
    # td2fc of the source pulse
    ref_sf = otr.td2fc(opt.s_tpsf[None, :], f, t, 1)[0]
    assert rel(opt.s_fpsf, ref_sf) < 1e-13
    td = tr.fc2td(torch.as_tensor(ref), f, t[:300], 1).cpu().numpy()
    assert rel(td, otr.fc2td(ref, f, t[:300], 1)) < 1e-12
    gates = np.stack([((t >= a) & (t < a + 500e-12)).astype(float) for a in (0.0, 200e-12, 400e-12)])
    gfc = otr.td2fc(gates, f, t, 1)
    gfc_gpu = tr.td2fc(torch.as_tensor(gates), f, t, 1).cpu().numpy()
    assert rel(gfc_gpu, gfc) < 1e-13
    tg = tr.fc2tg(torch.as_tensor(ref), gfc, f, t.size, 1).cpu().numpy()
    assert rel(tg, otr.fc2tg(ref, gfc, f, t.size, 1)) < 1e-12


def test_jacobian_tg(setup):
    from dogpup_amd import _lib
    from oracle import jacobian as ojac, optode as oopt, solver as osol, transforms as otr
    D, A, om, opt = setup["D"], setup["A"], setup["om"], setup["opt"]
    t, f, nf = setup["t"], setup["f"], setup["nf"]
    sid, sw = opt.s_bary
    did, dw = opt.d_bary
    phi_ref = osol.direct(A, oopt.source_matrix(om, sid, sw, opt.s_fpsf))
    phiA_ref = osol.direct(A, oopt.source_matrix(om, did, dw, opt.s_fpsf))
    gates = np.stack([((t >= a) & (t < a + 500e-12)).astype(float) for a in (0.0, 200e-12, 400e-12)])
    gfc = otr.td2fc(gates, f, t, 1)
    link = setup["link"]
    dx = 2.0
    ref = ojac.j_tg(phi_ref, phiA_ref, link, opt.s_fpsf, dx, gfc, f, t.size)   # NM x NG x NV
    fw = D.solve_point_sources("sd", mode="fp64")
    ns, nd = len(sid), len(did)
    w = np.where(np.arange(nf) == 0, 0.5, 1.0)
    coef = -2.0 * t.size * w[None, :] * np.conj(gfc) * dx ** 3 / opt.s_fpsf[None, :]
    ng, nm, nn = gfc.shape[0], link.shape[0], D.nn
    # forward and adjoint live in one block: columns 0..ns-1 and ns..ns+nd-1
    phi_s = torch.view_as_real(fw.data)
    out = torch.empty((nn, ng, nm), dtype=torch.float32, device="cuda")
    link_all = np.stack([link[:, 0], link[:, 1] + ns], 1)
    _lib.call("dogpup_jacobian_tg", _lib.ptr(phi_s),
              _lib.ptr(torch.view_as_real(dev(coef, torch.complex128))), _lib.ptr(dev(link_all, torch.int32)),
              None, _lib.ptr(out), nn, nf, ng, ns + nd, fw.nsp, nm, nn, _lib.stream_ptr())
    got = out.cpu().numpy()                      # [v][g][m]
    assert rel(np.transpose(got, (2, 1, 0)), ref) < 1e-5
    # the package wrapper gives the same tensor
    from dogpup_amd.jacobian import jacobian_tg
    J2, info = jacobian_tg(D, torch.as_tensor(gfc), mode="fp64", tol=1e-12, dxyz=dx, field=fw)
    print("coef diff", np.abs(info["coef"].cpu().numpy() - coef).max(), "link diff",
          np.abs(info["link"].cpu().numpy() - link_all).max(), "J diff", (J2 - out).abs().max().item(),
          "shapes", tuple(J2.shape), tuple(out.shape))
    assert rel(np.transpose(J2.cpu().numpy(), (2, 1, 0)), ref) < 1e-5


def test_fc2td_field_mfma(setup):
    """fc2td over a solved nodal field on FP64 MFMA against the FP64 oracle
    fc2td(phi, f, t, 2) of the same field, for 300 time bins (two 128-bin
    passes and a tail): every curve to the FP32 rounding of its peak."""
    from dogpup_amd import transforms as tr
    from oracle import optode as oopt, solver as osol, transforms as otr
    D, A, om, opt = setup["D"], setup["A"], setup["om"], setup["opt"]
    sid, sw = opt.s_bary
    phi = osol.direct(A, oopt.source_matrix(om, sid, sw, opt.s_fpsf))       # NN x NF x NS
    t = setup["t"][:300]
    ref = otr.fc2td(phi, setup["f"], t, 1)                                     # NN x NT x NS
    fw = D.solve_point_sources("s", mode="fp64")
    print("[case] fc2td_field", flush=True)
    td = tr.fc2td_field(fw, setup["f"], t)
    got = tr.td_field_to_reference(td, fw.ncols).cpu().numpy()
    assert got.shape == ref.shape
    assert rel(got, ref) < 1e-5
    same = otr.fc2td(fw.to_reference_layout().cpu().numpy(), setup["f"], t, 1)
    curve = np.abs(got - same).max(axis=1) / np.abs(same).max(axis=1)
    assert curve.max() < 1e-6, curve.max()


@pytest.mark.parametrize("ng,nf", [(11, 7), (3, 64)])
def test_jacobian_tg_mfma_vs_scalar_and_fp64(monkeypatch, ng, nf):
    """The MFMA J-TG kernel (link block as a 32x32 real GEMM per node and
    gate) against the scalar kernel and an FP64 einsum, on random fields with
    a shuffled partial link list, several column blocks, an odd frequency
    count, gate blocks of 8 and a Rytov-type scale."""
    _need_gpu()
    from dogpup_amd import _lib
    rng = np.random.default_rng(5)
    nn, nsp, ncols = 1000, 8, 40
    ncb = ncols // nsp
    field = rng.standard_normal((nf * ncb, nn, nsp)) + 1j * rng.standard_normal((nf * ncb, nn, nsp))
    F = field.reshape(nf, ncb, nn, nsp).transpose(0, 1, 3, 2).reshape(nf, ncols, nn)   # [k][col][v]
    scol = rng.permutation(ncols)[:20]
    dcol = rng.permutation(ncols)[:25]
    pairs = np.array([(a, b) for a in scol for b in dcol])
    link = pairs[rng.permutation(len(pairs))[:300]].astype(np.int32)
    nm = link.shape[0]
    coef = rng.standard_normal((ng, nf)) + 1j * rng.standard_normal((ng, nf))
    scale = rng.uniform(0.5, 2.0, (ng, nm)).astype(np.float32)
    ref = np.einsum("gk,kmv->vgm", coef, F[:, link[:, 0], :] * F[:, link[:, 1], :]).real * scale[None]
    print(f"[case] jtg mfma ng {ng} nf {nf}", flush=True)
    outs = {}
    for flag in ("0", "1"):
        if flag == "1":
            monkeypatch.setenv("DOGPUP_JTG_SCALAR", "1")
        out = torch.empty((nn, ng, nm), dtype=torch.float32, device="cuda")
        _lib.call("dogpup_jacobian_tg", _lib.ptr(torch.view_as_real(dev(field, torch.complex128))),
                  _lib.ptr(torch.view_as_real(dev(coef, torch.complex128))), _lib.ptr(dev(link, torch.int32)),
                  _lib.ptr(dev(scale, torch.float32)), _lib.ptr(out), nn, nf, ng, ncols, nsp, nm, nn,
                  _lib.stream_ptr())
        outs[flag] = out.cpu().numpy()
    monkeypatch.delenv("DOGPUP_JTG_SCALAR")
    assert rel(outs["0"], ref) < 1e-6
    assert rel(outs["1"], ref) < 1e-6
    assert rel(outs["0"], outs["1"]) < 1e-6


def test_jacobian_tg_cancellation(monkeypatch):
    """A J entry that is a near-total cancellation of its frequency terms, as
    for early gates and distant pairs: unit-modulus fields, coefficients
    C_gk = w_gk conj(phi_s phi_d) with terms w_gk of order 1 whose sum is
    1e-6 (times the gate number + 1).  Both kernels sum in FP64, so J comes
    out to the FP32 rounding of the stored value; FP32 sums would leave an
    error of order 1e-5 here, ten times the value."""
    _need_gpu()
    import math
    from dogpup_amd import _lib
    rng = np.random.default_rng(11)
    nn, nsp, ncols, nf, ng = 64, 8, 2, 64, 3
    a, b = rng.uniform(0, 2 * np.pi, nf), rng.uniform(0, 2 * np.pi, nf)
    field = np.zeros((nf, nn, nsp), complex)
    field[:, :, 0] = np.exp(1j * a)[:, None]
    field[:, :, 1] = np.exp(1j * b)[:, None]
    P = np.exp(1j * (a + b))
    w = rng.standard_normal((ng, nf))
    w -= w.mean(axis=1, keepdims=True)
    w += (1e-6 * (np.arange(ng) + 1) / nf)[:, None]
    truth = np.array([math.fsum(w[g] * np.abs(P) ** 2) for g in range(ng)])
    coef = w * np.conj(P)[None, :]
    link = np.array([[0, 1]], np.int32)
    print("[case] jtg cancellation", flush=True)
    for flag in ("0", "1"):
        monkeypatch.setenv("DOGPUP_JTG_SCALAR", flag)
        out = torch.empty((nn, ng, 1), dtype=torch.float32, device="cuda")
        _lib.call("dogpup_jacobian_tg", _lib.ptr(torch.view_as_real(dev(field, torch.complex128))),
                  _lib.ptr(torch.view_as_real(dev(coef, torch.complex128))), _lib.ptr(dev(link, torch.int32)),
                  None, _lib.ptr(out), nn, nf, ng, ncols, nsp, 1, nn, _lib.stream_ptr())
        got = out.cpu().numpy()[:, :, 0].astype(np.float64)
        assert np.all(np.abs(got - truth[None, :]) <= 1e-6 * np.abs(truth)[None, :]), (flag, got[0], truth)


def test_autograd_vjp_vs_fd_and_reference_jacobian(setup):
    """AbsorptionTG: exact-mode gradient vs central finite differences in
    nodal mua (kappa fixed); nodal-mode gradient at dxyz = 2 vs the
    contraction of the oracle's reference Jacobian at dxyz = 2; the mesh is
    left as it was (same properties, bit-identical system); the nodal mode
    needs a dxyz and refuses a voxel basis."""
    from dogpup_amd import jacobian as jac
    from oracle import jacobian as ojac, optode as oopt, solver as osol, transforms as otr
    D, om, opt = setup["D"], setup["om"], setup["opt"]
    t, f, nf = setup["t"], setup["f"], setup["nf"]
    gates = np.stack([((t >= a) & (t < a + 500e-12)).astype(float) for a in (0.0, 200e-12, 400e-12)])
    gfc = otr.td2fc(gates, f, t, 1)
    mua_before, kappa_before = D.mua.copy(), D.kappa.copy()
    K_before, M_before = D.K.clone(), D.M.clone()
    mua0 = torch.tensor(D.mua.copy(), dtype=torch.float64, requires_grad=True)
    rng = np.random.default_rng(5)
    ybar = torch.as_tensor(rng.normal(size=(opt.link.shape[0], gfc.shape[0])), device="cuda")
    y = jac.absorption_tg(mua0, D, gfc, True, "fp64", 1e-13)
    (y * ybar).sum().backward()
    g_exact = mua0.grad.numpy().copy()
    nodes = [int(np.argmin(np.linalg.norm(D.node - np.array(p), axis=1))) for p in [(10, 10, 6), (7, 12, 8), (14, 9, 3)]]
    for v in nodes:
        d = 1e-5
        vals = []
        for sgn in (1, -1):
            mu = mua0.detach().numpy().copy()
            mu[v] += sgn * d
            with torch.no_grad():
                yy = jac.absorption_tg(torch.as_tensor(mu), D, gfc, True, "fp64", 1e-13)
            vals.append(float((yy * ybar).sum()))
        fd = (vals[0] - vals[1]) / (2 * d)
        assert abs(fd - g_exact[v]) <= 1e-5 * abs(g_exact).max(), (v, fd, g_exact[v])
    # the forward calls restored the mesh
    assert np.array_equal(D.mua, mua_before) and np.array_equal(D.kappa, kappa_before)
    assert torch.equal(D.K, K_before) and torch.equal(D.M, M_before)
    # nodal mode, dxyz = 2, against the oracle's J^T ybar
    assert getattr(D, "basis", None) is None
    saved = getattr(D, "dxyz", None)
    D.dxyz = None
    mu1 = torch.tensor(D.mua.copy(), dtype=torch.float64, requires_grad=True)
    with pytest.raises(ValueError, match="dxyz"):
        jac.absorption_tg(mu1, D, gfc, False, "fp64", 1e-13).sum().backward()
    mu1 = torch.tensor(D.mua.copy(), dtype=torch.float64, requires_grad=True)
    y = jac.absorption_tg(mu1, D, gfc, False, "fp64", 1e-13, dxyz=2.0)
    (y * ybar).sum().backward()
    g_nodal = mu1.grad.numpy()
    sid, sw = opt.s_bary
    did, dw = opt.d_bary
    phi = osol.direct(setup["A"], oopt.source_matrix(om, sid, sw, opt.s_fpsf))
    phiA = osol.direct(setup["A"], oopt.source_matrix(om, did, dw, opt.s_fpsf))
    Jref = ojac.j_tg(phi, phiA, setup["link"], opt.s_fpsf, 2.0, gfc, f, t.size)   # NM x NG x NV
    gref = np.einsum("mg,mgv->v", ybar.cpu().numpy(), Jref)
    assert rel(g_nodal, gref) < 1e-8
    # and it is jacobian_tg's J^T ybar at the same dxyz
    Jt, _ = jac.jacobian_tg(D, torch.as_tensor(gfc), mode="fp64", tol=1e-13, dxyz=2.0)
    gJ = torch.einsum("vgm,mg->v", Jt.double(), ybar).cpu().numpy()
    assert rel(g_nodal, gJ) < 1e-5                                  # J_TG is float32
    D.dxyz = saved


@pytest.mark.parametrize("method", ["bicgstab", "cocg"])
def test_graph_replay_matches_eager(setup, monkeypatch, method):
    """The captured-iteration path (HIP graph replayed between convergence
    polls, several solves on one handle) gives exactly the eager result."""
    from dogpup_amd import mesh as dm
    D = setup["D"]
    res = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("DOGPUP_GRAPH", flag)
        print(f"[case] graph {flag} {method}", flush=True)
        S = dm.FluenceSolver(D, 4, dm.MODES["mixed"])
        S.set_method(method)
        nodes, w = D._bary_dev("s")
        spec = torch.as_tensor(D.optode.s_fpsf, dtype=torch.complex128, device="cuda")
        b = S.empty_block()
        from dogpup_amd import _lib
        _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), D.nn, S.nf, 4, S.nsp,
                  _lib.ptr(nodes), _lib.ptr(w), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
        outs = []
        for rep in range(3):
            f = S.solve(b, tol=1e-10, maxit=500)
            torch.cuda.synchronize()
            outs.append(f.data.clone())
            assert (f.status == 0).all()
        for o in outs[1:]:
            assert torch.equal(o, outs[0])
        res[flag] = outs[0]
    assert torch.equal(res["0"], res["1"])


@pytest.mark.parametrize("nsp", [16, 32])
def test_halo_spmv_matches_gather_spmv(setup, monkeypatch, nsp):
    """The complex64 SpMVs' x-halo form (chunk halos staged in LDS) and the
    gather form give the same mixed COCG solve: iteration counts within a few
    steps (they round differently), fields equal at the tolerance, both
    within tolerance of the direct solve."""
    from dogpup_amd import mesh as dm, _lib
    from oracle import optode as oopt, solver as osol
    D, A, om, opt = setup["D"], setup["A"], setup["om"], setup["opt"]
    sid, sw = opt.s_bary
    ref = osol.direct(A, oopt.source_matrix(om, sid, sw, opt.s_fpsf))
    out = {}
    for halo in ("1", "0"):
        monkeypatch.setenv("DOGPUP_HALO", halo)
        print(f"[case] halo {halo} nsp {nsp}", flush=True)
        S = dm.FluenceSolver(D, len(sid), dm.MODES["mixed"], nsp=nsp)
        S.set_method("cocg")
        nodes, w = D._bary_dev("s")
        spec = torch.as_tensor(opt.s_fpsf, dtype=torch.complex128, device="cuda")
        b = S.empty_block()
        _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), D.nn, S.nf, len(sid), S.nsp,
                  _lib.ptr(nodes), _lib.ptr(w), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
        f = S.solve(b, tol=1e-10)
        assert (f.status == 0).all()
        out[halo] = (f.to_reference_layout().cpu().numpy(), f.iters.copy())
        del S
    assert np.abs(out["1"][1] - out["0"][1]).max() <= 3
    assert rel(out["1"][0], out["0"][0]) < 1e-8
    assert rel(out["1"][0], ref) < 1e-8 and rel(out["0"][0], ref) < 1e-8


@pytest.mark.parametrize("method", ["cocg", "bicgstab"])
def test_refinement_zero_rhs(setup, method):
    """Refinement form with zero right-hand sides (the reference's relative
    residual divides by their zero initial residual, dMesh.m:1285; here they
    are answered x = 0 with DOGPUP_RHS_ZERO): a zero column comes back
    exactly 0 while the others match the direct solve; an all-zero block
    ends the solve before any sweep and still writes x = 0 over the caller's
    buffer (filled with NaN beforehand)."""
    from dogpup_amd import mesh as dm, _lib
    from oracle import optode as oopt, solver as osol
    D, A, om, opt = setup["D"], setup["A"], setup["om"], setup["opt"]
    sid, sw = opt.s_bary
    ref = osol.direct(A, oopt.source_matrix(om, sid, sw, opt.s_fpsf))
    print(f"[case] zero right-hand sides, mixed {method}", flush=True)
    S = dm.FluenceSolver(D, len(sid), dm.MODES["mixed"])
    S.set_method(method)
    nodes, w = D._bary_dev("s")
    spec = torch.as_tensor(opt.s_fpsf, dtype=torch.complex128, device="cuda")
    b = S.empty_block()
    _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), D.nn, S.nf, len(sid), S.nsp,
              _lib.ptr(nodes), _lib.ptr(w), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
    assert S.ncb == 1                               # column s of every group is source s
    b[:, :, 1] = 0
    x = torch.full_like(b, float("nan"))
    f = S.solve(b, tol=1e-10, x=x)
    got = f.to_reference_layout().cpu().numpy()
    assert (f.status[:, 1] == _lib.RHS_ZERO).all() and (f.status[:, [0, 2, 3]] == 0).all(), f.status
    assert np.all(got[:, :, 1] == 0)
    for s_ in (0, 2, 3):
        assert rel(got[:, :, s_], ref[:, :, s_]) < 1e-7, s_
    x0 = torch.full_like(b, float("nan"))
    f0 = S.solve(torch.zeros_like(b), tol=1e-10, x=x0)
    assert (f0.status == _lib.RHS_ZERO).all()
    assert torch.count_nonzero(x0).item() == 0


@pytest.mark.parametrize("mode,tol", [("mixed", 1e-10), ("fp64", 1e-12)])
def test_adjoint_tolerance_per_column(setup, mode, tol):
    """Forward and adjoint batched in one solve with a looser tolerance on the
    detector columns (dogpup_solver_set_rhs_tol): each column meets its own
    tolerance, the detector columns stop earlier, the source columns still
    match the direct solve to the forward tolerance, and the setting is
    cleared by the next solve without one.  The reference-structure form
    (one lockstep tolerance) rejects it."""
    from dogpup_amd import _lib
    from oracle import optode as oopt, solver as osol
    D, A, om, opt = setup["D"], setup["A"], setup["om"], setup["opt"]
    sid, sw = opt.s_bary
    did, dw = opt.d_bary
    ns = len(sid)
    ref_s = osol.direct(A, oopt.source_matrix(om, sid, sw, opt.s_fpsf))
    ref_d = osol.direct(A, oopt.source_matrix(om, did, dw, opt.s_fpsf))
    adj_tol = 1e-5
    print(f"[case] per-column tolerance, {mode} cocg, tol {tol}, adjoint {adj_tol}", flush=True)
    f = D.solve_point_sources("sd", mode=mode, tol=tol, method="cocg", adj_tol=adj_tol)
    assert (f.status == 0).all(), f.status
    assert (f.relres[:, :ns] < tol).all(), f.relres
    assert (f.relres[:, ns:] < adj_tol).all(), f.relres
    assert f.relres[:, ns:].max() > 10 * tol
    assert f.iters[:, ns:].mean() < f.iters[:, :ns].mean(), f.iters
    got = f.to_reference_layout().cpu().numpy()
    bound_s = 1e-7 if mode == "mixed" else 1e-9
    for k in range(setup["nf"]):
        for s_ in range(ns):
            assert rel(got[:, k, s_], ref_s[:, k, s_]) < bound_s, (k, s_)
        for d_ in range(len(did)):
            assert rel(got[:, k, ns + d_], ref_d[:, k, d_]) < 100 * adj_tol, (k, d_)
    f2 = D.solve_point_sources("sd", mode=mode, tol=tol, method="cocg")
    assert (f2.status == 0).all() and (f2.relres < tol).all(), f2.relres
    S = D.solver(ns + len(did), "plain")
    with pytest.raises(_lib.DogpupError):
        S.set_col_tol(np.full(ns + len(did), 1e-6))


def test_assembly_true_robin_faces_switch():
    """The Robin-face switch (reference_faces=False: the true boundary faces
    instead of the reference's selection, dMesh.m:1129-1140): on a jittered
    slab, 1'K1 = sum_e vol_e * mean(mua) + R * (box surface) exactly, 1'M1 =
    box volume, and the reference selection adds R times the area of its
    extra faces."""
    from dogpup_amd import mesh as dm, topology
    m = small_slab(seed=4, inclusions=[((10.0, 10.0, 5.0), 3.0, 0.03)])
    props = np.stack([m.mua, m.musp], 1)
    print("[case] true Robin faces", flush=True)
    Dt = dm.DMesh(m.node, m.elem, m.nr, props, reference_faces=False)
    Dr = dm.DMesh(m.node, m.elem, m.nr, props, reference_faces=True)
    ex, ey, ez = m.extent
    e = m.elem
    p = m.node[e]
    vol = np.abs(np.einsum("ij,ij->i", np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), p[:, 3] - p[:, 0])) / 6
    int_mua = float((vol * m.mua[e].mean(axis=1)).sum())
    R = topology.robin_factor(m.nr)
    surface = 2 * (ex * ey + ey * ez + ex * ez)
    Kt = float(Dt.K.sum())
    assert abs(Kt - (int_mua + R * surface)) <= 1e-12 * Kt, (Kt, int_mua + R * surface)
    assert abs(float(Dt.M.sum()) - ex * ey * ez) <= 1e-12 * ex * ey * ez
    bf = Dr.bface
    pf = m.node[bf]
    a = 0.5 * np.linalg.norm(np.cross(pf[:, 1] - pf[:, 0], pf[:, 2] - pf[:, 0]), axis=1).sum()
    assert a > surface                                   # the reference selection has extra faces
    Kr = float(Dr.K.sum())
    assert abs(Kr - (int_mua + R * a)) <= 1e-12 * Kr, (Kr, int_mua + R * a)
