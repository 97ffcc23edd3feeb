"""GPU parity of the LM dense algebra (MFMA GEMM, FP64 Cholesky) vs numpy and
the oracle."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def test_mfma_gemm_layout_exact():
    """Small-integer operands: FP32 products and sums are exact, so any lane
    layout mistake shows as an exact mismatch."""
    from dogpup_amd import recon
    rng = np.random.default_rng(0)
    for M, N, K in [(16, 16, 4), (70, 33, 29), (130, 65, 100)]:
        A = rng.integers(-4, 5, size=(M, K)).astype(np.float32)
        B = rng.integers(-4, 5, size=(N, K)).astype(np.float32)
        for exact in (False, True):
            C = recon.gemm_nt(A, B, exact=exact).cpu().numpy()
            assert np.array_equal(C, (A.astype(np.float64) @ B.T.astype(np.float64))), exact
        Bn = rng.integers(-4, 5, size=(K, N)).astype(np.float32)
        C2 = recon.gemm_nn(A, Bn).cpu().numpy()
        assert np.array_equal(C2, (A @ Bn))
        Cl = recon.gemm_nt(A, A, lower_only=True).cpu().numpy()
        ref = A.astype(np.float64) @ A.T.astype(np.float64)
        assert np.array_equal(np.tril(Cl), np.tril(ref))


def test_mfma_gemm_split_k_exact():
    """Shapes whose K is split over several blocks (few output tiles over a
    long K: 9 slices for 70 x 33 x 20000, 4 for the lower half of 130 x 130
    x 9000) with small-integer operands: every partial and total is exact,
    so the slices' sum equals the product exactly, in both precisions and in
    the lower-only form."""
    from dogpup_amd import recon
    rng = np.random.default_rng(4)
    for M, N, K in [(70, 33, 20000), (130, 130, 9000)]:
        A = rng.integers(-4, 5, size=(M, K)).astype(np.float32)
        B = rng.integers(-4, 5, size=(N, K)).astype(np.float32)
        ref = A.astype(np.float64) @ B.T.astype(np.float64)
        for exact in (False, True):
            C = recon.gemm_nt(A, B, exact=exact).cpu().numpy()
            assert np.array_equal(C, ref), (M, N, K, exact)
        Bn = np.ascontiguousarray(B.T)
        assert np.array_equal(recon.gemm_nn(A, Bn).cpu().numpy(), (A.astype(np.float64) @ Bn).astype(np.float32))
        Cl = recon.gemm_nt(A, A, lower_only=True).cpu().numpy()
        refl = A.astype(np.float64) @ A.T.astype(np.float64)
        assert np.array_equal(np.tril(Cl), np.tril(refl))


def test_cholesky_solve():
    from dogpup_amd import recon
    rng = np.random.default_rng(1)
    for n in [5, 32, 33, 100, 300]:
        X = rng.normal(size=(n, n))
        A = X @ X.T + n * np.eye(n)
        L = torch.as_tensor(np.tril(A), device="cuda").contiguous()
        recon.chol_factor(L, 0.5)
        Lh = np.tril(L.cpu().numpy())
        assert np.allclose(Lh @ Lh.T, A + 0.5 * np.eye(n), rtol=1e-12, atol=1e-10)
        for nrhs in (3, 243):
            B = rng.normal(size=(n, nrhs))
            Xs = recon.chol_solve(L, B).cpu().numpy()
            assert np.allclose((A + 0.5 * np.eye(n)) @ Xs, B, rtol=1e-9, atol=1e-9)


def test_tik_inv_and_reshape_J_vs_oracle():
    from dogpup_amd import recon
    from oracle import recon as orec
    rng = np.random.default_rng(2)
    J = rng.normal(size=(40, 500)).astype(np.float32)
    y = rng.normal(size=40)
    got = recon.tik_inv(torch.as_tensor(J), 0.05, y).cpu().numpy()
    assert rel(got, orec.tik_inv(J.astype(np.float64), 0.05, y)) < 1e-4
    Js = rng.normal(size=(7, 500)).astype(np.float32)
    w = recon.reshape_J(torch.as_tensor(J), torch.as_tensor(Js), 1e-3).cpu().numpy()
    assert rel(w, orec.reshape_J(J.astype(np.float64), Js.astype(np.float64), 1e-3)) < 1e-4
