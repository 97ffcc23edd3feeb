"""ctypes driver for the fake-mex harness: NumPy arrays in and out of
mxArrays (column-major, interleaved complex), and calls of a gateway's
mexFunction that turn its MATLAB errors into MexError."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "lib"
DOUBLE, INT32 = 6, 12


class MexError(Exception):
    def __init__(self, ident: str, msg: str):
        super().__init__(f"{ident}: {msg}")
        self.ident, self.msg = ident, msg


class Mex:
    def __init__(self):
        self.fm = C.CDLL(str(LIB / "libfakemex.so"))
        vp = C.c_void_p
        self.fm.fm_create.restype = vp
        self.fm.fm_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_size_t), vp]
        self.fm.fm_data.restype = vp
        self.fm.fm_data.argtypes = [vp]
        self.fm.mxGetNumberOfDimensions.restype = C.c_size_t
        self.fm.mxGetNumberOfDimensions.argtypes = [vp]
        self.fm.mxGetDimensions.restype = C.POINTER(C.c_size_t)
        self.fm.mxGetDimensions.argtypes = [vp]
        self.fm.mxGetClassID.restype = C.c_int
        self.fm.mxGetClassID.argtypes = [vp]
        self.fm.mxIsComplex.restype = C.c_bool
        self.fm.mxIsComplex.argtypes = [vp]
        self.fm.mxDestroyArray.argtypes = [vp]
        self.fm.fm_invoke.restype = C.c_int
        self.fm.fm_invoke.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp), C.c_char_p, C.c_char_p,
                                      C.c_int]
        self._gw = {}

    def array(self, a) -> int:
        """an mxArray copy of a (a 1-D array becomes a column vector, a
        Python int an int32 scalar, a float a double scalar)"""
        if isinstance(a, int):
            a = np.array([[a]], dtype=np.int32)
        elif isinstance(a, float):
            a = np.array([[a]])
        a = np.asarray(a)
        if a.ndim == 1:
            a = a[:, None]
        if a.dtype == np.int32:
            cls, cplx = INT32, 0
        elif a.dtype == np.complex128:
            cls, cplx = DOUBLE, 1
        elif a.dtype == np.float64:
            cls, cplx = DOUBLE, 0
        else:
            raise TypeError(a.dtype)
        buf = np.asfortranarray(a)
        dims = (C.c_size_t * a.ndim)(*a.shape)
        return self.fm.fm_create(cls, cplx, a.ndim, dims, buf.ctypes.data if buf.size else None)

    def to_numpy(self, p: int) -> np.ndarray:
        nd = self.fm.mxGetNumberOfDimensions(p)
        dims = tuple(self.fm.mxGetDimensions(p)[i] for i in range(nd))
        cls, cplx = self.fm.mxGetClassID(p), bool(self.fm.mxIsComplex(p))
        dt = np.int32 if cls == INT32 else (np.complex128 if cplx else np.float64)
        n = int(np.prod(dims))
        out = np.zeros(dims, dtype=dt, order="F")
        if n:
            src = self.fm.fm_data(p)
            C.memmove(out.ctypes.data, src, n * out.itemsize)
        return out

    def gateway(self, name: str, variant: str = "rec"):
        so = LIB / (f"{name}_rec.so" if variant == "rec" else f"{name}.so")
        if so not in self._gw:
            self._gw[so] = C.CDLL(str(so))
        return self._gw[so]

    def call(self, name: str, *args, nlhs: int = 1, variant: str = "rec"):
        """[outputs] = name(args...); raises MexError for a MATLAB error"""
        lib = self.gateway(name, variant)
        fn = C.cast(lib.mexFunction, C.c_void_p)
        ins = [self.array(a) for a in args]
        prhs = (C.c_void_p * max(len(ins), 1))(*ins)
        plhs = (C.c_void_p * max(nlhs, 1))()
        ident, msg = C.create_string_buffer(512), C.create_string_buffer(512)
        try:
            rc = self.fm.fm_invoke(fn, nlhs, plhs, len(ins), prhs, ident, msg, 512)
            if rc:
                raise MexError(ident.value.decode(), msg.value.decode())
            outs = [self.to_numpy(plhs[i]) for i in range(nlhs)]
        finally:
            for p in ins:
                self.fm.mxDestroyArray(p)
            for i in range(nlhs):
                if plhs[i]:
                    self.fm.mxDestroyArray(plhs[i])
        return outs[0] if nlhs == 1 else outs


def recorder():
    """controls of the recording stand-in that the *_rec gateways link"""
    rec = C.CDLL(str(LIB / "librecord_dogpup.so"))
    rec.rec_fail.argtypes = [C.c_int, C.c_char_p]

    def last():
        fn, ints, sums, calls = C.c_int(), (C.c_int * 4)(), (C.c_double * 6)(), C.c_int()
        rec.rec_get(C.byref(fn), ints, sums, C.byref(calls))
        return fn.value, list(ints), list(sums), calls.value
    return rec, last


def checksum(a) -> float:
    """the stand-in's position-weighted checksum of an array's storage
    (interleaved re/im for complex)"""
    v = np.asarray(a)
    v = v.ravel(order="F")
    if np.iscomplexobj(v):
        v = np.stack([v.real, v.imag], 1).ravel()
    w = (np.arange(v.size) % 7) + 1
    return float((v.astype(np.float64) * w).sum())
