"""Dense algebra of the weighted LM reconstruction on the GPU.

``tik_inv`` and ``reshape_J`` mirror ``DOGPUP/reconstruction/tik_inv.m`` and
``reshape_J.m``.  J is rows x NV (float32 on the device); the normal matrix
J J^T (rows x rows) is formed on MFMA (by default exact FP64 products of the
FP32 entries with FP64 accumulation, because the reference's damping of
1e-6 * max(diag) is below FP32 rounding; FP32 MFMA on request) and
factorised in FP64.  tik_inv is solved in the data space,
    (J^T J + mu I)^-1 J^T y = J^T (J J^T + mu I)^-1 y,   mu = lam * max_j sum_i J_ij^2,
which is the same update without the NV x NV matrix (580 GB at 269k voxels).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

DEV = "cuda"


def _f32(a):
    t = a if torch.is_tensor(a) else torch.as_tensor(np.asarray(a))
    return t.to(DEV, torch.float32).contiguous()


def _f64(a):
    t = a if torch.is_tensor(a) else torch.as_tensor(np.asarray(a))
    return t.to(DEV, torch.float64).contiguous()


def gemm_nt(A, B, out_f64=True, lower_only=False, alpha=1.0, exact=True):
    """C = alpha * A B^T from FP32 operands; exact=True: FP64 MFMA with FP64
    accumulation, exact=False: FP32 MFMA with FP32 accumulation."""
    A, B = _f32(A), _f32(B)
    M, K = A.shape
    N = B.shape[0]
    assert B.shape[1] == K
    C = torch.zeros((M, N), dtype=torch.float64 if out_f64 else torch.float32, device=DEV)
    _lib.call("dogpup_gemm_f32", _lib.ptr(A), _lib.ptr(B), _lib.ptr(C), M, N, K, K, K, N, float(alpha), 1,
              int(out_f64), int(lower_only), int(exact), _lib.stream_ptr())
    return C


def gemm_nn(A, B, alpha=1.0, exact=False):
    """C = alpha * A B (FP32 output; FP32 MFMA, or FP64 MFMA when exact)."""
    A, B = _f32(A), _f32(B)
    M, K = A.shape
    N = B.shape[1]
    assert B.shape[0] == K
    C = torch.empty((M, N), dtype=torch.float32, device=DEV)
    _lib.call("dogpup_gemm_f32", _lib.ptr(A), _lib.ptr(B), _lib.ptr(C), M, N, K, K, N, N, float(alpha), 0, 0,
              0, int(exact), _lib.stream_ptr())
    return C


def col_sumsq(J):
    J = _f32(J)
    out = torch.empty(J.shape[1], dtype=torch.float64, device=DEV)
    _lib.call("dogpup_col_sumsq", _lib.ptr(J), J.shape[0], J.shape[1], _lib.ptr(out), _lib.stream_ptr())
    return out


def gemv_t(J, y):
    J = _f32(J)
    y = _f64(y)
    out = torch.empty(J.shape[1], dtype=torch.float64, device=DEV)
    _lib.call("dogpup_gemv_t", _lib.ptr(J), _lib.ptr(y), J.shape[0], J.shape[1], _lib.ptr(out), _lib.stream_ptr())
    return out


def chol_factor(C, shift=0.0):
    """In-place lower Cholesky of (C + shift I) (FP64)."""
    info = np.zeros(1, np.int32)
    _lib.call("dogpup_chol_factor", _lib.ptr(C), C.shape[0], float(shift), _lib.ptr(info), _lib.stream_ptr())
    return C


def chol_solve(L, B):
    """Solve (L L^T) X = B in place (B: n or n x nrhs, FP64)."""
    B = _f64(B).clone()
    nrhs = 1 if B.dim() == 1 else B.shape[1]
    _lib.call("dogpup_chol_solve", _lib.ptr(L), _lib.ptr(B), L.shape[0], nrhs, _lib.stream_ptr())
    return B


def tik_inv(J, lam, y):
    """update = (J'J + lam*alpha*I) \\ (J'y), alpha = max column sum of J.^2
    (tik_inv.m:11-19), via the data-space identity."""
    J = _f32(J)
    alpha = float(col_sumsq(J).max())
    C = gemm_nt(J, J, out_f64=True, lower_only=True)
    chol_factor(C, lam * alpha)
    z = chol_solve(C, y)
    return gemv_t(J, z)


def reshape_J(J, Js, lam):
    """w = (Js J') / (J J' + lam*max(diag(J J'))*I) (reshape_J.m:13-15)."""
    J = _f32(J)
    C = gemm_nt(J, J, out_f64=True, lower_only=True)
    dmax = float(torch.diagonal(C).max())
    chol_factor(C, lam * dmax)
    S = gemm_nt(_f32(Js), J, out_f64=True)            # NW x rows
    X = chol_solve(C, S.t().contiguous())            # rows x NW
    return X.t().contiguous()
