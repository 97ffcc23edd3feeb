"""Oracle: FSAI preconditioner pattern and factors (test infrastructure only).

Restates ``DOGPUP/dMesh.m:1351-1385`` (FSAIP_gen: pattern selection on
tril(A)), ``dMesh.m:1388-1414`` (sparse_csr) and the per-row dense solve of
``cuda_lib/mex_sources/FSAIP_CUDA.cu:28-160`` (gather the symmetric local
matrix, partial-pivot complex Gaussian elimination against e_m, scale by the
principal square root of the last entry).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def pattern_rows_cols(A_tril: sp.spmatrix, N: int = 30):
    """Pattern (row, col) pairs of G chosen by dMesh.m:1361-1375.

    MATLAB ``find`` enumerates column-major; ``sortrows([i |v|],'descend')`` is a
    stable sort on (row desc, |v| desc); the kept positions are each row
    group's first entry and the N-1 positions before it in that order."""
    C = sp.csc_matrix(A_tril)
    C.sort_indices()
    coo = C.tocoo()
    # column-major enumeration as MATLAB find: csc order is (col, row)
    i = coo.row.astype(np.int64)
    j = coo.col.astype(np.int64)
    v = np.abs(coo.data)
    order = np.lexsort((-v, -i))          # primary: row desc, secondary |v| desc, stable
    rows_sorted = i[order]
    first = np.concatenate([[rows_sorted[-1] + 1], rows_sorted])
    d = np.diff(first)
    starts = np.flatnonzero(d) + 1        # 1-based positions as in MATLAB
    cand = (starts[None, :] - np.arange(N)[:, None]).ravel()
    cand = np.unique(cand)
    cand = cand[cand > 0]
    keep = order[cand - 1]
    pr, pc = i[keep], j[keep]
    o = np.lexsort((pc, pr))
    return pr[o], pc[o]


def csr_from_pairs(r, c, n):
    """sparse_csr(r,c,v) (dMesh.m:1397-1412): sorted 0-based CSR pointers."""
    o = np.lexsort((c, r))
    r, c = r[o], c[o]
    rptr = np.zeros(n + 1, np.int64)
    np.add.at(rptr, r + 1, 1)
    return np.cumsum(rptr).astype(np.int32), c.astype(np.int32)


def complex_sqrt(z):
    """FSAIP_CUDA.cu:28-43 (principal branch, b==0 special case)."""
    a, b = z.real, z.imag
    mod = np.sqrt(a * a + b * b)
    re = np.sqrt((mod + a) / 2)
    if b == 0:
        if a < 0:
            return complex(0.0, np.sqrt(-a))
        return complex(re, 0.0)
    return complex(re, (1 if b > 0 else -1) * np.sqrt((mod - a) / 2))


def solve_pivot(Aloc, rhs):
    """Partial-pivot Gaussian elimination then back substitution
    (FSAIP_CUDA.cu:45-90)."""
    n = Aloc.shape[0]
    M = np.concatenate([Aloc.astype(complex), rhs.astype(complex)[:, None]], axis=1)
    for i in range(n):
        piv = i
        for j in range(i + 1, n):
            if abs(M[j, i]) > abs(M[piv, i]):
                piv = j
        if piv != i:
            M[[i, piv], i:] = M[[piv, i], i:]
        for j in range(i + 1, n):
            fac = M[j, i] / M[i, i]
            M[j, i:] -= fac * M[i, i:]
    x = np.zeros(n, complex)
    for i in range(n - 1, -1, -1):
        s = M[i, i + 1:n] @ x[i + 1:n]
        x[i] = (M[i, n] - s) / M[i, i]
    return x


def fsai_row(Ak: sp.csr_matrix, cols):
    """G row for pattern ``cols`` (ascending, diagonal last) of A_k."""
    m = len(cols)
    Aloc = np.asarray(Ak[cols][:, cols].todense())
    e = np.zeros(m)
    e[-1] = 1.0
    x = solve_pivot(Aloc, e)
    return x / complex_sqrt(x[-1])


def fsaip_gen(A_full: sp.csr_matrix, nf: int, N: int = 30):
    """Pattern of G (CSR) and values valG (nnzG x NF), frequency k using
    A_k = Re(A) + i*k*Im(A)."""
    n = A_full.shape[0]
    A_tril = sp.tril(A_full).tocsr()
    pr, pc = pattern_rows_cols(A_tril, N)
    rptr, cptr = csr_from_pairs(pr, pc, n)
    valG = np.zeros((cptr.size, nf), complex)
    for k in range(nf):
        Ak = (A_full.real + 1j * k * A_full.imag).tocsr()
        for i in range(n):
            s, e = rptr[i], rptr[i + 1]
            valG[s:e, k] = fsai_row(Ak, cptr[s:e])
    return rptr, cptr, valG


def transpose_csr(rptr, cptr, vals, n):
    """CSR of G^T with values permuted (dMesh.m:1255-1256)."""
    G = sp.csr_matrix((np.arange(cptr.size) + 1, cptr, rptr), shape=(n, n))
    GT = G.T.tocsr()
    GT.sort_indices()
    perm = GT.data.astype(np.int64) - 1
    return GT.indptr.astype(np.int32), GT.indices.astype(np.int32), vals[perm]


def g_matrix(rptr, cptr, vals_k, n):
    return sp.csr_matrix((vals_k, cptr, rptr), shape=(n, n))


def fsaip_gen_batched(A_full: sp.csr_matrix, nf: int, N: int = 30, fscale=None):
    """Same factors as ``fsaip_gen`` with rows of equal pattern length solved
    together by LAPACK (partial-pivot LU), for meshes where the per-row
    restatement is too slow.  Agrees with ``fsai_row`` to roundoff
    (tests/test_oracle_fsai.py)."""
    n = A_full.shape[0]
    A_tril = sp.tril(A_full).tocsr()
    pr, pc = pattern_rows_cols(A_tril, N)
    rptr, cptr = csr_from_pairs(pr, pc, n)
    lens = np.diff(rptr)
    valG = np.zeros((cptr.size, nf), complex)
    Are = sp.csr_matrix(A_full.real)
    Aim = sp.csr_matrix(A_full.imag)
    for m in np.unique(lens):
        rows = np.flatnonzero(lens == m)
        idx = rptr[rows][:, None] + np.arange(m)[None, :]
        P = cptr[idx]                                        # R x m
        ii = np.repeat(P[:, :, None], m, axis=2)
        jj = np.repeat(P[:, None, :], m, axis=1)
        re = np.asarray(Are[ii.ravel(), jj.ravel()]).reshape(ii.shape)
        im = np.asarray(Aim[ii.ravel(), jj.ravel()]).reshape(ii.shape)
        e = np.zeros((len(rows), m, 1))
        e[:, -1, 0] = 1.0
        for k in range(nf):
            fk = k if fscale is None else fscale[k]
            x = np.linalg.solve(re + 1j * fk * im, e)[:, :, 0]
            q = np.sqrt(x[:, -1])                            # principal branch
            valG[idx.ravel(), k] = (x / q[:, None]).ravel()
    return rptr, cptr, valG
