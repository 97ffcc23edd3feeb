"""CPU tests of the Galerkin guess weights of the frequency continuation
(dogpup_amd/csrc/rb_weights.hpp), built with g++ into a small test library.

The weights combine converged solutions x_j of A_j x = b_j, A_j = K + i f_j M,
b_j = s_j q, into the Galerkin solution of A_k x = b_k in their span, from
the M-Gram matrix x_i^H M x_j and the dots x_i^H q alone.  Pinned by:
* a span that is the whole space (more solutions than unknowns): the guess is
  the exact solution;
* a smaller span: the guess equals the Galerkin projection computed directly
  from A_k with an orthonormal basis (QR), which also checks the identity
  x_i^H K x_j = s_j x_i^H q - i f_j x_i^H M x_j the weights rely on;
* a swapped orientation of the Gram matrix or of the scale difference fails
  the exactness test.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "rbw", "lib", "librbw.so")


@pytest.fixture(scope="module")
def rbw():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "dogpup_amd", "csrc"),
                    "-o", LIB, os.path.join(HERE, "rbw", "rbw_driver.cpp")], check=True)
    lib = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.rbw.argtypes = [ctypes.c_int, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]

    def call(G, cq, sj, fj, sk, fk):
        n = len(cq)
        G = np.ascontiguousarray(G, complex)
        cq = np.ascontiguousarray(cq, complex)
        sj = np.ascontiguousarray(sj, complex)
        fj = np.ascontiguousarray(fj, float)
        y = np.zeros(n, complex)
        p = lambda a: a.ctypes.data_as(dp)  # noqa: E731
        lib.rbw(n, p(G), p(cq), p(sj), p(fj), float(np.real(sk)), float(np.imag(sk)), float(fk), p(y))
        return y
    return call


def problem(nn, freqs, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((nn, nn))
    K = B @ B.T + nn * np.eye(nn)
    C = rng.standard_normal((nn, nn))
    M = C @ C.T + np.eye(nn)
    q = rng.standard_normal(nn) + 1j * rng.standard_normal(nn)
    s = np.exp(-0.3j * np.asarray(freqs)) * (1.0 + 0.1 * np.asarray(freqs))
    X = np.stack([np.linalg.solve(K + 1j * f * M, sf * q) for f, sf in zip(freqs, s)], axis=1)
    return K, M, q, s, X


def weights(rbw, M, q, s, freqs, X, sk, fk):
    G = X.conj().T @ M @ X
    return rbw(G, X.conj().T @ q, s, freqs, sk, fk)


def test_whole_space_span_gives_the_exact_solution(rbw):
    freqs = np.array([0.0, 1.0, 2.0, 4.0, 5.0, 7.0])
    K, M, q, s, X = problem(4, freqs, 0)
    for fk in (3.0, 6.0, 0.5):
        sk = np.exp(-0.3j * fk) * (1.0 + 0.1 * fk)
        y = weights(rbw, M, q, s, freqs, X, sk, fk)
        exact = np.linalg.solve(K + 1j * fk * M, sk * q)
        assert np.linalg.norm(X @ y - exact) <= 1e-9 * np.linalg.norm(exact)


def test_subspace_guess_is_the_galerkin_projection(rbw):
    freqs = np.array([0.0, 2.0, 4.0, 6.0, 8.0, 10.0])
    K, M, q, s, X = problem(40, freqs, 1)
    for fk in (3.0, 5.0, 9.0):
        sk = np.exp(-0.3j * fk) * (1.0 + 0.1 * fk)
        y = weights(rbw, M, q, s, freqs, X, sk, fk)
        Q, _ = np.linalg.qr(X)
        Ak = K + 1j * fk * M
        z = np.linalg.solve(Q.conj().T @ Ak @ Q, Q.conj().T @ (sk * q))
        ref = Q @ z
        assert np.linalg.norm(X @ y - ref) <= 1e-8 * np.linalg.norm(ref)
        # and it is a good guess: far better than the nearest solution alone
        exact = np.linalg.solve(Ak, sk * q)
        near = X[:, np.argmin(np.abs(freqs - fk))] * sk / s[np.argmin(np.abs(freqs - fk))]
        assert np.linalg.norm(X @ y - exact) < 0.1 * np.linalg.norm(near - exact)


def test_planted_mistakes_fail_exactness(rbw):
    freqs = np.array([0.0, 1.0, 2.0, 4.0, 5.0, 7.0])
    K, M, q, s, X = problem(4, freqs, 2)
    fk = 3.0
    sk = np.exp(-0.3j * fk) * (1.0 + 0.1 * fk)
    exact = np.linalg.solve(K + 1j * fk * M, sk * q)
    G = X.conj().T @ M @ X
    cq = X.conj().T @ q
    for bad in (rbw(G.T, cq, s, freqs, sk, fk),                 # Gram transposed (conjugated)
                rbw(G, cq.conj(), s, freqs, sk, fk),                 # x^T q for x^H q
                rbw(G, cq, s, freqs, sk, -fk)):                 # sign of the scale difference
        assert np.linalg.norm(X @ bad - exact) > 1e-3 * np.linalg.norm(exact)


def test_degenerate_inputs_give_zero_or_reduced_weights(rbw):
    # a zero window (zero right-hand side): no guess
    y = rbw(np.zeros((3, 3)), np.zeros(3), np.ones(3), np.array([0.0, 1.0, 2.0]), 1.0, 1.5)
    assert np.all(y == 0)
    # repeated solutions: the pivoting keeps one copy and the guess is exact
    freqs = np.array([0.0, 1.0, 2.0, 3.0])
    K, M, q, s, X = problem(3, freqs, 3)
    Xr = np.concatenate([X, X[:, :2]], axis=1)
    fr = np.concatenate([freqs, freqs[:2]])
    sr = np.concatenate([s, s[:2]])
    fk = 1.5
    sk = np.exp(-0.3j * fk) * (1.0 + 0.1 * fk)
    y = weights(rbw, M, q, sr, fr, Xr, sk, fk)
    exact = np.linalg.solve(K + 1j * fk * M, sk * q)
    assert np.all(np.isfinite(y))
    assert np.linalg.norm(Xr @ y - exact) <= 1e-9 * np.linalg.norm(exact)
