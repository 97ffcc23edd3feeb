"""Pins the LM dense-algebra oracle (oracle/recon.py)."""
import numpy as np

from oracle import recon


def test_tik_inv_push_through_identity():
    rng = np.random.default_rng(0)
    J = rng.normal(size=(12, 40))
    y = rng.normal(size=12)
    lam = 0.3
    mu = lam * (J ** 2).sum(axis=0).max()
    data_space = J.T @ np.linalg.solve(J @ J.T + mu * np.eye(12), y)
    assert np.allclose(recon.tik_inv(J, lam, y), data_space, rtol=1e-10, atol=1e-12)


def test_tik_inv_orthogonal_columns_closed_form():
    # orthogonal columns: H is diagonal, update_j = (J_j . y) / (|J_j|^2 + lam*alpha)
    Q, _ = np.linalg.qr(np.random.default_rng(1).normal(size=(10, 4)))
    J = Q * np.array([1.0, 2.0, 0.5, 3.0])
    y = np.arange(10.0)
    lam = 0.1
    alpha = 9.0
    expect = (J.T @ y) / (np.array([1.0, 4.0, 0.25, 9.0]) + lam * alpha)
    assert np.allclose(recon.tik_inv(J, lam, y), expect)


def test_reshape_J_limits():
    rng = np.random.default_rng(2)
    J = rng.normal(size=(6, 30))
    # Js = J and vanishing damping: w -> I
    w = recon.reshape_J(J, J, 1e-12)
    assert np.allclose(w, np.eye(6), atol=1e-8)
    # w (J J' + mu I) = Js J'
    Js = rng.normal(size=(3, 30))
    lam = 0.2
    w = recon.reshape_J(J, Js, lam)
    A = J @ J.T
    assert np.allclose(w @ (A + lam * np.diag(A).max() * np.eye(6)), Js @ J.T)
