"""Oracle: FSAI-preconditioned BiCGSTAB and the direct solve (test infrastructure only).

``bicgstab_fsai`` restates ``DOGPUP/dMesh.m:1247-1348`` (gpuBicstab_FSAIP):
right preconditioning with M^-1 = G^T G, unconjugated bilinear forms, the
reciprocal step ``alph``, the G-weighted omega, x0 = 0 and lockstep convergence
on the maximum relative residual over all columns.  ``direct`` is the
"backslash" oracle (sparse LU per frequency).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import fsai as ofsai


def _apply(mats, X):
    # X: NN x NS x NF ; mats: list of NF sparse matrices
    return np.stack([mats[k] @ X[:, :, k] for k in range(X.shape[2])], axis=2)


def bicgstab_fsai(A_full: sp.csr_matrix, Q, tol=1e-12, maxit=1000, fsai=None):
    """Q: NN x NF x NS complex.  Returns (x NN x NF x NS, iterations, relres)."""
    n, nf, ns = Q.shape
    if fsai is None:
        fsai = ofsai.fsaip_gen(A_full, nf)
    rptr, cptr, valG = fsai
    Gs = [ofsai.g_matrix(rptr, cptr, valG[:, k], n) for k in range(nf)]
    GTs = [G.T.tocsr() for G in Gs]
    As = [(A_full.real + 1j * k * A_full.imag).tocsr() for k in range(nf)]
    b = np.transpose(Q, (0, 2, 1)).astype(complex)           # nodes x sources x freqs
    x, it, err = bicgstab_core(As, Gs, GTs, b, tol, maxit)
    return np.transpose(x, (0, 2, 1)), it, err


def bicgstab_core(As, Gs, GTs, b, tol=1e-12, maxit=1000):
    """The iteration of dMesh.m:1283-1341 on given matrices A_k, G_k, G_k^T
    (lists over frequencies) for b (NN x NS x NF); returns (x, iterations,
    largest relative residual)."""
    x = np.zeros_like(b)
    r = b - _apply(As, x)
    err0 = np.sqrt((np.abs(r) ** 2).sum(axis=0))
    rhat0 = r.copy()
    rho0 = (rhat0 * r).sum(axis=0)
    p = r.copy()
    it = 0
    err = np.inf
    for it in range(1, maxit + 1):
        h = _apply(GTs, _apply(Gs, p))
        v = _apply(As, h)
        alph = (v * rhat0).sum(axis=0) / rho0
        h = x + h / alph
        s = r - v / alph
        err = (np.sqrt((np.abs(s) ** 2).sum(axis=0)) / err0).max()
        if err < tol:
            x = h
            break
        Gsv = _apply(Gs, s)
        z = _apply(GTs, Gsv)
        t = _apply(As, z)
        Gtt = _apply(Gs, t)
        om = (Gtt * Gsv).sum(axis=0) / (Gtt * Gtt).sum(axis=0)
        x = h + om * z
        r = s - om * t
        p = p - om * v
        err = (np.sqrt((np.abs(r) ** 2).sum(axis=0)) / err0).max()
        if err < tol:
            break
        rho = (rhat0 * r).sum(axis=0)
        beta = rho / rho0 * (1.0 / (alph * om))
        rho0 = rho
        p = r + beta * p
    return x, it, err


def direct(A_full: sp.csr_matrix, Q):
    """x_k = A_k \\ Q(:,k,:) by sparse LU (the backslash oracle)."""
    n, nf, ns = Q.shape
    x = np.zeros((n, nf, ns), complex)
    for k in range(nf):
        Ak = (A_full.real + 1j * k * A_full.imag).tocsc()
        lu = spla.splu(Ak)
        x[:, k, :] = lu.solve(np.ascontiguousarray(Q[:, k, :].astype(complex)))
    return x


def smv(A_full, f, X):
    """smv_mex semantics: y(:,c,k) = (Re A + i f[k] Im A) x(:,c,k)
    (smv_CUDA.cu:8-54), x NN x NS x NF."""
    out = np.zeros_like(X, dtype=complex)
    for k in range(X.shape[2]):
        Ak = (A_full.real + 1j * f[k] * A_full.imag).tocsr()
        out[:, :, k] = Ak @ X[:, :, k]
    return out


def fsaimv(rptr, cptr, valG, X):
    """FSAImv_mex semantics: y(:,c,k) = G_k x(:,c,k) (FSAImv_CUDA.cu:8-53)."""
    n = X.shape[0]
    out = np.zeros_like(X, dtype=complex)
    for k in range(X.shape[2]):
        out[:, :, k] = ofsai.g_matrix(rptr, cptr, valG[:, k], n) @ X[:, :, k]
    return out
