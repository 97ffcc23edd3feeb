"""ctypes binding of libdogpup_hip.so (the C ABI declared in include/dogpup.h).

The library is loaded from dogpup_amd/lib/ (built in-tree by
``python -m dogpup_amd.build``).  There is no fallback: if the library is
missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libdogpup_hip.so"

OK = 0
ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_NOMEM = -1, -2, -3, -4
RHS_CONVERGED, RHS_MAXIT, RHS_BREAKDOWN, RHS_ZERO = 0, 1, 2, 3
MODE_C128, MODE_C64_REFINE, MODE_C128_PLAIN = 0, 1, 2
METHOD_BICGSTAB, METHOD_COCG = 0, 1
CONT_GALERKIN, CONT_LAGRANGE = 1, 2
CONT_KINDS = {"galerkin": CONT_GALERKIN, "lagrange": CONT_LAGRANGE}


def cont_kind(cont) -> int:
    """Frequency-continuation argument (True / None: Galerkin guesses, False:
    none, 'galerkin' or 'lagrange') as dogpup_solver_set_freq_continuation's
    `on`."""
    if cont is None or cont is True:
        return CONT_GALERKIN
    if cont is False:
        return 0
    if cont not in CONT_KINDS:
        raise ValueError(f"unknown continuation {cont!r}; expected True, False, 'galerkin' or 'lagrange'")
    return CONT_KINDS[cont]

_lock = threading.Lock()
_lib = None

vp, ip, dp, i32, f64 = C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_double
P = C.POINTER

# name -> argtypes (all return int)
SIGNATURES = {
    "dogpup_last_error": None,
    "dogpup_version": [],
    "dogpup_device_count": [P(C.c_int)],
    "dogpup_smv_call": [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, vp],
    "dogpup_fsaimv_call": [vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp],
    "dogpup_fsaip_call": [vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp],
    "dogpup_smv_mex": [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32],
    "dogpup_fsaimv_mex": [vp, vp, vp, vp, vp, i32, i32, i32, i32],
    "dogpup_fsaip_gen_mex": [vp, vp, vp, vp, vp, vp, i32, i32, i32, i32],
    "dogpup_assemble": [i32, i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, f64, vp, vp, vp, vp],
    "dogpup_solver_create": [P(vp), i32, i32, i32, i32, i32, i32, vp, vp, i32, vp, vp, vp],
    "dogpup_batch_create": [P(vp), i32, vp, vp, vp, i32, i32, i32, vp, vp, i32, vp, vp, vp],
    "dogpup_solver_set_tuning": [vp, i32, i32, i32, i32],
    "dogpup_solver_set_cont_stride": [vp, i32],
    "dogpup_solver_destroy": [vp],
    "dogpup_solver_set_matrix": [vp, vp, vp, vp, vp],
    "dogpup_solver_solve": [vp, vp, vp, f64, i32, vp, vp, vp, vp],
    "dogpup_solver_stats": [vp, vp, i32],
    "dogpup_solver_get_fsai": [vp, vp, vp],
    "dogpup_layout_to_solver": [vp, vp, i32, i32, i32, i32, vp],
    "dogpup_layout_from_solver": [vp, vp, i32, i32, i32, i32, vp],
    "dogpup_point_sources": [vp, i32, i32, i32, i32, vp, vp, vp, vp],
    "dogpup_td2fc": [vp, vp, vp, vp, i32, i32, i32, i32, vp],
    "dogpup_fc2td": [vp, vp, vp, vp, i32, i32, i32, vp],
    "dogpup_fc2td_field": [vp, vp, vp, vp, i32, i32, C.c_longlong, vp],
    "dogpup_fc2td_field_noalloc": [vp, vp, vp, vp, i32, i32, C.c_longlong, vp],
    "dogpup_fc2tg": [vp, vp, vp, i32, i32, i32, f64, vp],
    "dogpup_jacobian_tg": [vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, i32, vp],
    "dogpup_meas": [vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp],
    "dogpup_jacobian_vjp": [vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, i32, i32, f64, vp],
    "dogpup_gemm_f32": [vp, vp, vp, i32, i32, i32, C.c_longlong, C.c_longlong, C.c_longlong, C.c_float, i32, i32, i32, i32, vp],
    "dogpup_col_sumsq": [vp, i32, C.c_longlong, vp, vp],
    "dogpup_gemv_t": [vp, vp, i32, C.c_longlong, vp, vp],
    "dogpup_chol_factor": [vp, i32, f64, vp, vp],
    "dogpup_chol_solve": [vp, vp, i32, i32, vp],
    "dogpup_solver_set_profile": [vp, i32],
    "dogpup_solver_set_method": [vp, i32],
    "dogpup_solver_set_rhs_tol": [vp, vp, i32],
    "dogpup_solver_set_guess": [vp, i32, vp, vp],
    "dogpup_solver_set_freq_continuation": [vp, i32, vp],
    "dogpup_solver_profile": [vp, vp, i32],
    "dogpup_locator_create": [P(vp), vp, i32, vp, i32, vp],
    "dogpup_locator_destroy": [vp],
    "dogpup_locator_stats": [vp, vp],
    "dogpup_locator_find": [vp, vp, C.c_longlong, f64, vp, vp, vp, vp],
    "dogpup_grid_to_mesh": [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp],
    "dogpup_ell4_apply": [vp, vp, vp, vp, C.c_longlong, C.c_longlong, i32, i32, vp],
}


class DogpupError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dogpup error {code}: {msg}")
        self.code = code


def lib():
    """Load the library (once); raises if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(f"{LIB_PATH} is not built; run `python -m dogpup_amd.build`")
            L = C.CDLL(str(LIB_PATH))
            for name, args in SIGNATURES.items():
                fn = getattr(L, name, None)
                if fn is None:
                    continue
                if name == "dogpup_last_error":
                    fn.restype = C.c_char_p
                    fn.argtypes = []
                else:
                    fn.restype = C.c_int
                    fn.argtypes = args
            _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return [n for n in SIGNATURES if hasattr(L, n)]


def call(name, *args):
    """Call an entry point and raise DogpupError on a non-zero status."""
    L = lib()
    fn = getattr(L, name)
    rc = fn(*args)
    if rc != OK:
        raise DogpupError(rc, L.dogpup_last_error().decode(errors="replace"))
    return rc


class _Ptr(C.c_void_p):
    """void* that keeps the array it points into alive for as long as the
    argument tuple of the call holds it (temporaries stay valid during the
    call; device work queued on torch's current stream stays valid after it,
    by the caching allocator's stream ordering)."""


def ptr(t):
    """Device/host pointer of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        if not t.is_contiguous():
            raise ValueError("non-contiguous tensor passed to the C ABI")
        p = _Ptr(t.data_ptr())
    else:
        if not t.flags["C_CONTIGUOUS"]:
            raise ValueError("non-contiguous array passed to the C ABI")
        p = _Ptr(t.ctypes.data)
    p.keep = t
    return p


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
