"""The OpenMP FP64 restatement of the FSAI setup and the reference-structure
BiCGSTAB (oracle/cpu_bicgstab.cpp, the timed CPU baseline) against the
pinned NumPy oracle: the per-row partial-pivot factors (FSAIP_CUDA.cu:28-160)
to roundoff, the same BiCGSTAB iterate count and solution as the NumPy
restatement of dMesh.m:1247-1348, and the direct solve."""
import numpy as np
import pytest

from conftest import rel, small_slab


@pytest.fixture(scope="module")
def sys1():
    from dogpup_amd import meshgen
    from oracle import cpu, fem, fsai as ofsai, optode as oopt, transforms as otr
    cpu.lib()
    m = small_slab(seed=5, inclusions=[((10.0, 10.0, 5.0), 3.0, 0.03)])
    om = fem.make_mesh(m.node, m.elem, m.nr, m.mua, m.musp)
    t = np.arange(0, 501) * 20e-12
    f, df = otr.freq_axis(t, 4)
    A = fem.full_system(om, df)
    src = meshgen.top_grid(m.extent, 2, 2, 6.0, 6.0, 8.0, 8.0)
    snap = oopt.snap2mesh(om, src, src)
    Q = oopt.source_matrix(om, snap["s_id"], snap["s_w"], np.ones(4))
    return A, Q


def test_fsai_setup_matches_per_row_elimination(sys1):
    from oracle import cpu, fsai as ofsai
    A, _ = sys1
    rG, cG, valG = ofsai.fsaip_gen(A, 4)                 # frequencies k = 0..3 of A_k = Re A + i k Im A
    for k in (0, 3):
        Ak = (A.real + 1j * k * A.imag).tocsr()
        g = cpu.fsai_setup(Ak, rG, cG, nthreads=4)
        assert rel(g, valG[:, k]) < 1e-13


def test_bicgstab_matches_numpy_oracle_and_direct(sys1):
    from oracle import cpu, fsai as ofsai, solver as osol
    A, Q = sys1
    k = 2
    Ak = (A.real + 1j * k * A.imag).tocsr()
    rG, cG, valG = ofsai.fsaip_gen(A, 3)
    g = cpu.fsai_setup(Ak, rG, cG)
    b = Q[:, 0, :].astype(complex)
    x, it, rr = cpu.bicgstab(Ak, rG, cG, g, b, tol=1e-12, nthreads=3)
    # the NumPy restatement of the same iteration on A_k
    G = ofsai.g_matrix(rG, cG, valG[:, k], Ak.shape[0])
    xr, itr, err = osol.bicgstab_core([Ak], [G], [G.T.tocsr()], b[:, :, None], tol=1e-12)
    assert it == itr and rr < 1e-12
    assert rel(x, xr[:, :, 0]) < 1e-9
    import scipy.sparse.linalg as spla
    xd = spla.splu(Ak.tocsc()).solve(b)
    assert rel(x, xd) < 1e-9
