"""ctypes binding of oracle/lib/liboracle_cpu.so (test infrastructure
only): the FSAI setup (FSAIP_CUDA.cu:28-160) and the reference-structure
FSAI-BiCGSTAB (dMesh.m:1283-1341) for one frequency, FP64, on OpenMP host
threads."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import scipy.sparse as sp

LIB = Path(__file__).resolve().parent / "lib" / "liboracle_cpu.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built; run `python -m oracle.build_cpu`")
        L = C.CDLL(str(LIB))
        vp, i32 = C.c_void_p, C.c_int
        L.oracle_fsai_setup.argtypes = [i32, vp, vp, vp, vp, vp, vp, i32]
        L.oracle_fsai_setup.restype = i32
        L.oracle_bicgstab.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, C.c_double, i32, i32,
                                      C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.oracle_bicgstab.restype = i32
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def fsai_setup(A: sp.csr_matrix, rG, cG, nthreads=0):
    """FSAI factor values of the full symmetric complex A on the pattern
    (rG, cG) (ascending rows, diagonal last)."""
    A = A.tocsr()
    A.sort_indices()
    rA, cA = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    vA = np.ascontiguousarray(A.data, np.complex128)
    rG, cG = np.ascontiguousarray(rG, np.int32), np.ascontiguousarray(cG, np.int32)
    g = np.zeros(cG.size, np.complex128)
    if lib().oracle_fsai_setup(A.shape[0], _p(rA), _p(cA), _p(vA), _p(rG), _p(cG), _p(g), nthreads) != 0:
        raise ValueError("FSAI pattern row longer than 64 entries")
    return g


def bicgstab(A: sp.csr_matrix, rG, cG, g, b, tol=1e-12, maxit=1000, nthreads=0):
    """x (n x ns), iterations, final largest relative residual."""
    A = A.tocsr()
    A.sort_indices()
    n = A.shape[0]
    rA, cA = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    vA = np.ascontiguousarray(A.data, np.complex128)
    Gm = sp.csr_matrix((np.asarray(g, np.complex128), np.asarray(cG), np.asarray(rG)), shape=(n, n))
    GT = Gm.T.tocsr()
    GT.sort_indices()
    rG, cG = np.ascontiguousarray(rG, np.int32), np.ascontiguousarray(cG, np.int32)
    vG = np.ascontiguousarray(g, np.complex128)
    rGT, cGT = GT.indptr.astype(np.int32), GT.indices.astype(np.int32)
    vGT = np.ascontiguousarray(GT.data, np.complex128)
    b = np.ascontiguousarray(np.asarray(b, np.complex128).reshape(n, -1))
    x = np.zeros_like(b)
    it, rr = C.c_int(0), C.c_double(0.0)
    lib().oracle_bicgstab(n, _p(rA), _p(cA), _p(vA), _p(rG), _p(cG), _p(vG), _p(rGT), _p(cGT), _p(vGT), _p(b), _p(x),
                          b.shape[1], float(tol), int(maxit), nthreads, C.byref(it), C.byref(rr))
    return x, it.value, rr.value
