"""Oracle: dense algebra of the LM step (test infrastructure only).

Restates ``DOGPUP/reconstruction/tik_inv.m:1-20`` and ``reshape_J.m:1-16``
literally (primal normal equations, FP64).
"""
from __future__ import annotations

import numpy as np


def tik_inv(J, lam, y):
    """H = J'J + lam*max(sum(J.^2))*I; update = H \\ (J'y) (tik_inv.m:11-19)."""
    J = np.asarray(J, float)
    alph = (J ** 2).sum(axis=0).max()
    H = J.T @ J
    H[np.diag_indices_from(H)] += lam * alph
    return np.linalg.solve(H, J.T @ np.asarray(y, float))


def reshape_J(J, Js, lam):
    """A = J J'; A = A + lam*max(diag(A))*I; w = (Js J') / A (reshape_J.m:13-15)."""
    J = np.asarray(J, float)
    A = J @ J.T
    A = A + lam * np.diag(A).max() * np.eye(A.shape[0])
    return np.linalg.solve(A.T, (np.asarray(Js, float) @ J.T).T).T
