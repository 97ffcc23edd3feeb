"""The C ABI library loads and exports every entry point include/dogpup.h
declares (no GPU needed: nothing is launched)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_symbols():
    txt = (ROOT / "include" / "dogpup.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dogpup_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_reference_entry_points():
    syms = header_symbols()
    for s in ["dogpup_smv_call", "dogpup_fsaimv_call", "dogpup_fsaip_call", "dogpup_smv_mex",
              "dogpup_fsaimv_mex", "dogpup_fsaip_gen_mex", "dogpup_assemble", "dogpup_solver_solve",
              "dogpup_td2fc", "dogpup_fc2td", "dogpup_fc2tg", "dogpup_jacobian_tg"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    from dogpup_amd import _lib
    L = _lib.lib()
    missing = [s for s in header_symbols() if not hasattr(L, s)]
    assert not missing, missing
    for s in header_symbols():
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"


def test_argument_errors_without_gpu():
    """Argument validation happens before any device call."""
    import ctypes as C
    from dogpup_amd import _lib
    with pytest.raises(_lib.DogpupError) as e:
        _lib.call("dogpup_smv_call", None, None, None, None, None, None, None, None, None, 1, 1, 2, None)
    assert e.value.code == _lib.ERR_ARG
    h = C.c_void_p()
    with pytest.raises(_lib.DogpupError) as e:
        _lib.call("dogpup_solver_create", C.byref(h), 10, 2, 4, 12, 0, 1, None, None, 1, None, None, None)
    assert e.value.code == _lib.ERR_ARG and "group width" in str(e.value)
    with pytest.raises(_lib.DogpupError) as e:
        _lib.call("dogpup_solver_set_rhs_tol", None, None, 0)
    assert e.value.code == _lib.ERR_ARG and "NULL solver" in str(e.value)
