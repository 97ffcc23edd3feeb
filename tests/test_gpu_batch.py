"""Batches of uneven meshes in one solver handle (dogpup_batch_create).

A batch runs each group with the same arithmetic as a handle of its own mesh
(same group width, tiles and reductions), so its fields must equal the
single-mesh solves bit for bit; the FP64 batch must match the oracle's LU
solve of every mesh.
"""
import numpy as np
import pytest

from conftest import rel, small_slab

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _batch_meshes(morton=False):
    """Three uneven slabs; morton=True renumbers each along the solver's
    Morton curve first (the batch then iterates in the caller's x)."""
    from dogpup_amd import meshgen, mesh as dm
    t = np.arange(0, 501) * 20e-12
    specs = [((20.0, 20.0, 10.0), 2.0, 3, 4), ((14.0, 10.0, 8.0), 1.0, 5, 6), ((30.0, 24.0, 12.0), 3.0, 12, 3)]
    out = []
    for k, (ext, h, ns, nf) in enumerate(specs):
        m = small_slab(seed=k + 1, extent=ext, h=h)
        if morton:
            m = meshgen.sorted_mesh(m)
        nx = int(np.ceil(np.sqrt(ns)))
        ny = -(-ns // nx)
        src = meshgen.top_grid(m.extent, nx, ny, 0.25 * ext[0], 0.25 * ext[1], 0.5 * ext[0] / max(nx - 1, 1),
                               0.5 * ext[1] / max(ny - 1, 1))[:ns]
        det = meshgen.top_grid(m.extent, 2, 2, 0.3 * ext[0], 0.3 * ext[1], 0.4 * ext[0], 0.4 * ext[1])
        D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
        D.add_optode(dm.DOptode(src, det, meshgen.all_pairs(ns, 4), 1e3, meshgen.gaussian_tpsf(t, 250e-12), t, nf))
        out.append((D, m))
    return out


@pytest.mark.parametrize("mode,method,morton", [("mixed", "cocg", False), ("mixed", "bicgstab", False),
                                                ("fp64", "cocg", False), ("mixed", "cocg", True),
                                                ("fp64", "cocg", True)])
def test_batch_equals_single_mesh_solves(mode, method, morton):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import batch, mesh as dm
    meshes = [D for D, _ in _batch_meshes(morton)]
    print(f"[case] batch {mode} {method} morton={morton}", flush=True)
    B = batch.MeshBatch(meshes, "s", mode=mode, nsp=8, method=method)
    tol = 1e-10 if mode == "mixed" else 1e-12
    fields = B.solve(tol=tol)
    for D, f in zip(meshes, fields):
        assert (f.status == 0).all(), f.status
        S = dm.FluenceSolver(D, f.ncols, dm.MODES[mode], nsp=8)
        S.set_method(method)
        if mode == "mixed":
            S.set_continuation(True, D.optode.s_fpsf)          # as the batch does for each mesh
        nodes, w = D._bary_dev("s")
        spec = torch.as_tensor(D.optode.s_fpsf, dtype=torch.complex128, device="cuda")
        b = S.empty_block()
        from dogpup_amd import _lib
        _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), D.nn, S.nf, f.ncols, S.nsp,
                  _lib.ptr(nodes), _lib.ptr(w), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
        single = S.solve(b, tol=tol)
        assert torch.equal(single.data, f.data)
        assert (single.iters == f.iters).all()


@pytest.mark.parametrize("morton", [False, True])
def test_batch_fp64_vs_direct(morton):
    """Every mesh of an FP64 batch against the oracle's direct solve; with
    Morton-sorted meshes the handle's node order is the identity and the
    solve runs in the caller's b and x."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import batch
    from oracle import fem, optode as oopt, solver as osol, transforms as otr
    pairs = _batch_meshes(morton)
    if morton:
        assert all(np.array_equal(D.solver_order, np.arange(D.nn)) for D, _ in pairs)
    print(f"[case] fp64 batch vs direct, morton={morton}", flush=True)
    B = batch.MeshBatch([D for D, _ in pairs], "s", mode="fp64", nsp=8, method="cocg")
    fields = B.solve(tol=1e-12)
    for (D, m), f in zip(pairs, fields):
        om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
        _, df = otr.freq_axis(D.optode.tAxis, D.optode.Nf)
        A = fem.full_system(om, df)
        sid, sw = D.optode.s_bary
        ref = osol.direct(A, oopt.source_matrix(om, sid, sw, D.optode.s_fpsf))
        got = f.to_reference_layout().cpu().numpy()
        assert rel(got, ref) < 1e-9


def test_autotune_picks_a_measured_shape(tmp_path, monkeypatch):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import autotune
    monkeypatch.setenv("DOGPUP_TUNE_CACHE", str(tmp_path / "tune.json"))
    meshes = [D for D, _ in _batch_meshes()]
    best, table = autotune.tune(meshes, "s", "mixed", "cocg", iters=3, widths=(8, 16), tiles=(256, 512),
                                unrolls=(4, 8))
    assert len(table) == 8 and best in table
    assert best["s_per_iter"] == min(r["s_per_iter"] for r in table)
    again, _ = autotune.tune(meshes, "s", "mixed", "cocg")            # served from the cache
    assert again == best
    B = autotune.tuned_batch(meshes, "s", "mixed", "cocg")
    for f in B.solve(tol=1e-10):
        assert (f.status == 0).all()
    # the continuation spacing: whole solves of the chosen shape per stride,
    # the fastest kept and applied by tuned_batch; every spacing converges
    best2, table2 = autotune.tune(meshes, "s", "mixed", "cocg", iters=3, widths=(8,), tiles=(512,), unrolls=(8,),
                                  strides=(4, 8, 16), tol=1e-8)
    solves = [r for r in table2 if "s_per_solve" in r]
    assert [r["cont_stride"] for r in solves] == [4, 8, 16]
    assert best2["s_per_solve"] == min(r["s_per_solve"] for r in solves)
    assert best2["cont_stride"] == min(solves, key=lambda r: r["s_per_solve"])["cont_stride"]
    B2 = autotune.tuned_batch(meshes, "s", "mixed", "cocg", iters=3, widths=(8,), tiles=(512,), unrolls=(8,),
                              strides=(4, 8, 16), tol=1e-8)                  # served from the cache
    for f in B2.solve(tol=1e-10):
        assert (f.status == 0).all()


def test_split_batch_equals_single_mesh_handles():
    """A SplitBatch (parts of their own group width, one handle each) returns
    the fields in the caller's mesh order, each equal bit for bit to a
    handle of its mesh alone at its part's width; width_parts puts the
    meshes above the threshold in the wide part; parts must cover every mesh
    once."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import batch
    meshes = [D for D, _ in _batch_meshes()]                      # 3, 5 and 12 sources
    parts = batch.SplitBatch.width_parts([3, 5, 12], wide=16, narrow=8, threshold=8)
    assert parts == [([2], 16), ([0, 1], 8)]
    print("[case] split batch", flush=True)
    S = batch.SplitBatch(meshes, "s", "mixed", parts=parts, method="cocg")
    fields = S.solve(tol=1e-10)
    for i, (D, f) in enumerate(zip(meshes, fields)):
        nsp = 16 if i == 2 else 8
        assert f.nn == D.nn and f.nsp == nsp
        assert (f.status == 0).all(), f.status
        single = batch.MeshBatch([D], "s", "mixed", nsp=nsp, method="cocg").solve(tol=1e-10)[0]
        assert torch.equal(single.data, f.data)
        assert (single.iters == f.iters).all()
    with pytest.raises(ValueError):
        batch.SplitBatch(meshes, "s", "mixed", parts=[([0, 1], 8)])


def test_autotune_keeps_the_faster_of_single_handle_and_split(tmp_path, monkeypatch):
    """The tuner's partition search: whole solves of the single handle and of
    a SplitBatch partition, the faster kept ("split" only when the partition
    wins) and built by tuned_batch, whose solve converges."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dogpup_amd import autotune, batch
    monkeypatch.setenv("DOGPUP_TUNE_CACHE", str(tmp_path / "tune.json"))
    meshes = [D for D, _ in _batch_meshes()]
    kw = dict(iters=3, widths=(8,), tiles=(512,), unrolls=(8,), splits=((16, 8),), tol=1e-8)
    best, table = autotune.tune(meshes, "s", "mixed", "cocg", **kw)
    single = [r for r in table if "s_per_solve" in r and "split" not in r]
    split = [r for r in table if "split" in r]
    assert len(single) == 1 and len(split) == 1
    assert split[0]["parts"] == [[1, 16], [2, 8]]
    assert best["s_per_solve"] == min(single[0]["s_per_solve"], split[0]["s_per_solve"])
    assert ("split" in best) == (split[0]["s_per_solve"] < single[0]["s_per_solve"])
    B = autotune.tuned_batch(meshes, "s", "mixed", "cocg", **kw)          # served from the cache
    assert isinstance(B, batch.SplitBatch if "split" in best else batch.MeshBatch)
    for f in B.solve(tol=1e-10):
        assert (f.status == 0).all()
