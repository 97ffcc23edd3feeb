"""The complex64 SpMVs' x-halo form on meshes other than the box slab and on
tile shapes it does not take: same solves as the gather form and the
oracle's direct solve."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _mesh_and_ref(kind):
    from dogpup_amd import meshgen, mesh as dm
    from oracle import fem, optode as oopt, solver as osol, transforms as otr
    if kind == "cylinder":
        m = meshgen.cylinder(6.0, 5.0, 1.0)
    else:
        from conftest import small_slab
        m = small_slab(seed=9, extent=(16.0, 12.0, 8.0), h=1.0)
    c = m.extent[0] / 2.0
    src = np.stack([c + np.linspace(-2.5, 2.5, 10), np.full(10, m.extent[1] / 2.0), np.full(10, m.extent[2])], 1)
    det = np.array([[c, m.extent[1] / 2.0 + 1.5, m.extent[2]]])
    t = np.arange(0, 401) * 25e-12
    D = dm.DMesh(m.node, m.elem, m.nr, np.stack([m.mua, m.musp], 1))
    opt = dm.DOptode(src, det, [[0, 0]], 1e3, meshgen.gaussian_tpsf(t), t, 3)
    D.add_optode(opt)
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    _, df = otr.freq_axis(t, 3)
    sid, sw = opt.s_bary
    ref = osol.direct(fem.full_system(om, df), oopt.source_matrix(om, sid, sw, opt.s_fpsf))
    return D, opt, ref


def _solve(D, opt, tile=0, nsp=None):
    from dogpup_amd import mesh as dm, _lib
    sid, _ = opt.s_bary
    S = dm.FluenceSolver(D, len(sid), dm.MODES["mixed"], nsp=nsp)
    S.set_method("cocg")
    if tile:
        _lib.call("dogpup_solver_set_tuning", S.h, tile, 0, 0, 0)
    nodes, w = D._bary_dev("s")
    spec = torch.as_tensor(opt.s_fpsf, dtype=torch.complex128, device="cuda")
    b = S.empty_block()
    _lib.call("dogpup_point_sources", _lib.ptr(torch.view_as_real(b)), D.nn, S.nf, len(sid), S.nsp,
              _lib.ptr(nodes), _lib.ptr(w), _lib.ptr(torch.view_as_real(spec)), _lib.stream_ptr())
    f = S.solve(b, tol=1e-10)
    assert (f.status == 0).all(), f.status
    return f.to_reference_layout().cpu().numpy(), f.iters.copy()


@pytest.mark.parametrize("kind", ["cylinder", "slab"])
def test_halo_and_gather_forms_agree_with_direct(kind, monkeypatch):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    D, opt, ref = _mesh_and_ref(kind)
    print(f"[case] halo vs gather, {kind} {D.nn} nodes", flush=True)
    out = {}
    for halo in ("1", "0"):
        monkeypatch.setenv("DOGPUP_HALO", halo)
        out[halo] = _solve(D, opt)
    # the two forms round differently (one chunk's row sums are the same
    # sums in another order of fused multiply-adds), so iteration counts may
    # differ by a step or two; both converge to the same fluence
    assert np.abs(out["1"][1] - out["0"][1]).max() <= 3
    assert rel(out["1"][0], out["0"][0]) < 1e-8
    assert rel(out["1"][0], ref) < 1e-8 and rel(out["0"][0], ref) < 1e-8


def test_tiles_above_the_halo_limit_take_the_gather_form():
    """Tiles longer than the halo kernel's 1024 rows run the gather form
    (same solve)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    D, opt, ref = _mesh_and_ref("slab")
    print(f"[case] tile 2048 on {D.nn} nodes", flush=True)
    a, ia = _solve(D, opt)
    b, ib = _solve(D, opt, tile=2048)
    assert np.abs(ia - ib).max() <= 3
    assert rel(a, b) < 1e-8 and rel(b, ref) < 1e-8


@pytest.mark.parametrize("nsp", [8, 16, 32])
def test_interleaved_chunk_walk_matches_contiguous(nsp, monkeypatch):
    """With 128-row tiles on the 1,989-node slab every XCD holds two items of
    a group, which walk the range's four chunks interleaved (0, 2 and 1, 3)
    or contiguously (0, 1 and 2, 3); the halo SpMVs and the FP64 residual
    give the same solve either way (the per-item partial sums group the rows
    differently, so iteration counts may differ by a step or two)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    D, opt, ref = _mesh_and_ref("slab")
    assert D.nn >= 9 * 128            # at least two tiles per XCD
    out = {}
    for ilv in ("1", "0"):
        monkeypatch.setenv("DOGPUP_HALO_INTERLEAVE", ilv)
        print(f"[case] chunk walk interleave {ilv}, nsp {nsp}, tile 128, {D.nn} nodes", flush=True)
        out[ilv] = _solve(D, opt, tile=128, nsp=nsp)
    assert np.abs(out["1"][1] - out["0"][1]).max() <= 3
    assert rel(out["1"][0], out["0"][0]) < 1e-8
    assert rel(out["1"][0], ref) < 1e-8 and rel(out["0"][0], ref) < 1e-8
