"""Initial guesses of the refinement solver: frequency continuation (each
frequency of a later level started from the Galerkin combination, or the
Lagrange interpolation, of neighbouring frequencies solved before it) and
explicit guess terms (a warm start from the caller's x).  Every solution is checked against the oracle's
direct solve (dMesh.m:1238-1348 to convergence); the reference starts every
right-hand side from x0 = 0 (dMesh.m:1283), so the guesses may change the
iterations and nothing else."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel, small_slab

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sixteen():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import meshgen, mesh as dm
    from oracle import fem, optode as oopt, solver as osol, transforms as otr
    m = small_slab(seed=4, extent=(30.0, 30.0, 16.0), inclusions=[((15.0, 15.0, 8.0), 4.0, 0.03)])
    t = np.arange(0, 1001) * 20e-12
    nf = 16
    src = meshgen.top_grid(m.extent, 3, 3, 6.0, 6.0, 9.0, 9.0)
    det = meshgen.top_grid(m.extent, 2, 2, 10.0, 10.0, 10.0, 10.0)
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
    opt = dm.DOptode(src, det, meshgen.all_pairs(9, 4), 1e3, meshgen.gaussian_tpsf(t, 250e-12), t, nf)
    D.add_optode(opt)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    f, df = otr.freq_axis(t, nf)
    A = fem.full_system(om, df)
    sid, sw = opt.s_bary
    phi = osol.direct(A, oopt.source_matrix(om, sid, sw, opt.s_fpsf))
    return dict(D=D, phi=phi, m=m, src=src, det=det, t=t, nf=nf)


def _sorted_mesh_of(sixteen):
    """The fixture's mesh renumbered along the solver's Morton curve (a
    handle on it iterates in the caller's x) and the direct solution in that
    node order: the sources sit at the same points, so the fields are the
    same nodal values renumbered."""
    from dogpup_amd import meshgen, meshio, mesh as dm
    m = sixteen["m"]
    ms = meshgen.sorted_mesh(m)
    order = meshio.sortmesh(m.node, m.elem)[2]
    D = dm.DMesh(ms.node, ms.elem, ms.nr, np.stack([ms.mua, ms.musp], 1))
    D.add_optode(dm.DOptode(sixteen["src"], sixteen["det"], meshgen.all_pairs(9, 4), 1e3,
                            meshgen.gaussian_tpsf(sixteen["t"], 250e-12), sixteen["t"], sixteen["nf"]))
    assert np.array_equal(D.solver_order, np.arange(D.nn))
    return D, sixteen["phi"][order]


@pytest.mark.parametrize("method", ["cocg", "bicgstab"])
def test_frequency_continuation_vs_direct(sixteen, method):
    D, phi = sixteen["D"], sixteen["phi"]
    out = {}
    for cont in (False, "lagrange", "galerkin"):
        print(f"[case] continuation={cont} {method}", flush=True)
        fw = D.solve_point_sources("s", mode="mixed", tol=1e-10, method=method, cont=cont)
        assert (fw.status == 0).all(), fw.status
        assert (fw.relres < 1e-10).all()
        got = fw.to_reference_layout().cpu().numpy()
        for k in range(got.shape[1]):
            for s in range(got.shape[2]):
                assert rel(got[:, k, s], phi[:, k, s]) < 1e-7, (cont, k, s)
        out[cont] = fw.iters.copy()
    print("iterations per frequency (max over sources) without / Lagrange / Galerkin:",
          out[False].max(axis=1).tolist(), out["lagrange"].max(axis=1).tolist(),
          out["galerkin"].max(axis=1).tolist(), "totals", out[False].sum(), out["lagrange"].sum(),
          out["galerkin"].sum(), flush=True)
    # level 0 starts from x0 = 0 (Lagrange: every 8th frequency and the last;
    # Galerkin: every 16th and the last); every later level starts from a
    # guess and needs fewer iterations
    for cont, lvl0 in (("lagrange", [0, 8, 15]), ("galerkin", [0, 15])):
        assert (out[cont][lvl0] == out[False][lvl0]).all(), cont
        later = [k for k in range(16) if k not in lvl0]
        assert (out[cont][later].max(axis=1) < out[False][later].max(axis=1)).all(), cont
    # the Galerkin guesses (six solutions combined per column) beat the
    # Lagrange interpolation of four, with fewer frequencies solved from 0
    assert out["galerkin"].sum() < out["lagrange"].sum()


@pytest.mark.parametrize("mode", ["mixed", "fp64"])
def test_frequency_shards_vs_direct(sixteen, mode):
    """A handle over a subset of the optode's frequencies (one GPU's share when
    the frequencies are split across GPUs): two contiguous halves and a
    strided shard, each against the oracle's direct solve at its frequencies
    (dMesh.m:1238-1348), with the fields in the shard's frequency order."""
    D, phi = sixteen["D"], sixteen["phi"]
    for freqs in (np.arange(0, 8), np.arange(8, 16), np.array([13, 1, 9, 5])):
        print(f"[case] {mode} shard {freqs.tolist()}", flush=True)
        fw = D.solve_point_sources("s", mode=mode, tol=1e-10, freqs=freqs)
        assert fw.nf == len(freqs) and (fw.status == 0).all(), fw.status
        assert (fw.relres < 1e-10).all()
        got = fw.to_reference_layout().cpu().numpy()
        for j, k in enumerate(freqs):
            for s in range(got.shape[2]):
                assert rel(got[:, j, s], phi[:, k, s]) < 1e-7, (freqs.tolist(), k, s)
    with pytest.raises(ValueError):
        D.solve_point_sources("s", mode=mode, freqs=[0, sixteen["nf"]])


def test_warm_start_and_guess_terms(sixteen):
    """Explicit terms through the C ABI: a warm start from the previous
    solution converges at its first FP64 residual (0 iterations); a cycle
    and a term naming another mesh are rejected."""
    from dogpup_amd import _lib
    D, phi = sixteen["D"], sixteen["phi"]
    fw = D.solve_point_sources("s", mode="mixed", tol=1e-10, cont=True)
    S = D.solver(fw.ncols, "mixed")
    nf = S.nf
    src = np.arange(nf, dtype=np.int32)[:, None].copy()               # each frequency: itself
    w = np.zeros((nf, 1, 2))
    w[..., 0] = 1.0                                                    # complex weight 1
    _lib.call("dogpup_solver_set_guess", S.h, 1, _lib.ptr(src), _lib.ptr(w))
    S._cont_key = None
    b = S.empty_block(zero=False)
    spec = torch.as_tensor(D.optode.s_fpsf, dtype=torch.complex128, device="cuda")
    nodes, wq = D._bary_dev("s")
    _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), D.nn, nf, fw.ncols, S.nsp,
              _lib.ptr(nodes), _lib.ptr(wq), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
    x = fw.data.clone()
    f2 = S.solve(b, tol=1e-10, x=x)
    assert (f2.status == 0).all()
    assert f2.iters.max() == 0
    assert torch.equal(f2.data, fw.data)
    # a warm start from a perturbed solution converges to the same fields
    x = fw.data * (1.0 + 1e-3 * torch.randn_like(fw.data.real))
    f3 = S.solve(b, tol=1e-10, x=x)
    assert (f3.status == 0).all() and f3.iters.max() < fw.iters.max()
    got = f3.to_reference_layout().cpu().numpy()
    assert rel(got, phi) < 1e-7
    # rejected terms
    cyc = np.stack([np.roll(np.arange(nf), 1), np.full(nf, -1)], 1).astype(np.int32)
    _lib.call("dogpup_solver_set_guess", S.h, 2, _lib.ptr(cyc), _lib.ptr(np.ones((nf, 2, 2))))
    with pytest.raises(_lib.DogpupError) as e:
        S.solve(b, tol=1e-10)
    assert "cycle" in str(e.value)
    bad = np.full((nf, 1), nf, np.int32)
    with pytest.raises(_lib.DogpupError):
        _lib.call("dogpup_solver_set_guess", S.h, 1, _lib.ptr(bad), _lib.ptr(w))
    S.set_continuation(True, D.optode.s_fpsf)
    f4 = S.solve(b, tol=1e-10)
    assert (f4.status == 0).all()
    assert rel(f4.to_reference_layout().cpu().numpy(), phi) < 1e-7


def test_continuation_across_workspace_chunks(sixteen, monkeypatch):
    """A workspace budget that splits the 16 frequencies into three chunks:
    guesses read sources of earlier chunks from the caller's x, terms that
    name a later chunk are dropped, and every column still matches the
    direct solve."""
    from dogpup_amd import mesh as dm
    D0, phi = sixteen["D"], sixteen["phi"]
    monkeypatch.setenv("DOGPUP_WORKSPACE_GB", "0.015")       # about 6 of the 16 groups per chunk
    D = dm.DMesh(D0.node, D0.elem, D0.nr, np.stack([D0.mua, D0.musp], 1))
    D.optode = D0.optode                      # the same mesh: the placement carries over
    out = {}
    for cont in (False, True):
        print(f"[case] chunked, continuation={cont}", flush=True)
        fw = D.solve_point_sources("s", mode="mixed", tol=1e-10, cont=cont)
        assert (fw.status == 0).all()
        got = fw.to_reference_layout().cpu().numpy()
        for k in range(got.shape[1]):
            assert rel(got[:, k], phi[:, k]) < 1e-7, (cont, k)
        out[cont] = fw.iters.sum()
    print("iterations without / with continuation:", out[False], out[True], flush=True)
    assert out[True] < 0.8 * out[False]


def test_sorted_mesh_chunked_continuation_and_warm_start(sixteen, monkeypatch):
    """On a mesh in the solver's order the refinement form iterates in the
    caller's x: chunks of a small workspace read the earlier chunks'
    solutions from it and write their own in place, and a warm start reads
    each group's entry before its guess overwrites it.  Every column matches
    the direct solve, with and without continuation, and the warm start
    from a perturbed solution needs fewer iterations."""
    from dogpup_amd import _lib
    monkeypatch.setenv("DOGPUP_WORKSPACE_GB", "0.015")       # about 6 of the 16 groups per chunk
    D, phi = _sorted_mesh_of(sixteen)
    out = {}
    for cont in (False, True):
        print(f"[case] sorted mesh, chunked, continuation={cont}", flush=True)
        fw = D.solve_point_sources("s", mode="mixed", tol=1e-10, cont=cont)
        assert (fw.status == 0).all()
        got = fw.to_reference_layout().cpu().numpy()
        for k in range(got.shape[1]):
            assert rel(got[:, k], phi[:, k]) < 1e-7, (cont, k)
        out[cont] = fw
    assert out[True].iters.sum() < 0.8 * out[False].iters.sum()
    fw = out[True]
    S = D.solver(fw.ncols, "mixed")
    nf = S.nf
    src = np.arange(nf, dtype=np.int32)[:, None].copy()               # each frequency: itself
    w = np.zeros((nf, 1, 2))
    w[..., 0] = 1.0
    _lib.call("dogpup_solver_set_guess", S.h, 1, _lib.ptr(src), _lib.ptr(w))
    S._cont_key = None
    b = S.empty_block(zero=False)
    spec = torch.as_tensor(D.optode.s_fpsf, dtype=torch.complex128, device="cuda")
    nodes, wq = D._bary_dev("s")
    _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), D.nn, nf, fw.ncols, S.nsp,
              _lib.ptr(nodes), _lib.ptr(wq), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
    print("[case] sorted mesh, chunked warm start", flush=True)
    x = fw.data * (1.0 + 1e-3 * torch.randn_like(fw.data.real))
    f3 = S.solve(b, tol=1e-10, x=x)
    assert (f3.status == 0).all() and f3.iters.max() < fw.iters.max()
    assert rel(f3.to_reference_layout().cpu().numpy(), phi) < 1e-7
