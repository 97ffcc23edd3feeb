"""td2fc / fc2td / fc2tg and id_thresh on the GPU (mirrors of DOGPUP/td2fc.m,
fc2td.m, fc2tg.m, id_thresh.m).

Arrays are torch tensors (moved to ``cuda`` if they are not); the transformed
dimension ``dim`` is 0-based (the reference's ``dim`` is 1-based).  Results
are float64 / complex128 on the device.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

DEV = "cuda"


def _dev(x, dtype):
    t = torch.as_tensor(np.asarray(x) if not torch.is_tensor(x) else x)
    return t.to(device=DEV, dtype=dtype).contiguous()


def _front(x, dim):
    x = torch.movedim(x, dim, 0).contiguous()
    return x, x.shape


def td2fc(fun_td, f, t, dim):
    """F(k) = dt/P sum_t g(t) exp(-i w_k t), P = t(end) (td2fc.m:26-32)."""
    src = fun_td if torch.is_tensor(fun_td) else torch.as_tensor(np.asarray(fun_td))
    if dim >= src.dim():
        raise ValueError("Function dimensions and given dimension mismatch")
    f = _dev(f, torch.float64).reshape(-1)
    t = _dev(t, torch.float64).reshape(-1)
    if t.numel() != src.shape[dim]:
        raise ValueError("Size along given dimension does not match given time axis")
    cplx = src.is_complex()
    x, shp = _front(_dev(src, torch.complex128 if cplx else torch.float64), dim)
    ncol = int(np.prod(shp[1:], dtype=np.int64))
    out = torch.empty((f.numel(),) + tuple(shp[1:]), dtype=torch.complex128, device=DEV)
    xr = torch.view_as_real(x) if cplx else x
    _lib.call("dogpup_td2fc", _lib.ptr(xr), _lib.ptr(torch.view_as_real(out)), _lib.ptr(f), _lib.ptr(t),
              t.numel(), f.numel(), ncol, int(cplx), _lib.stream_ptr())
    return torch.movedim(out, 0, dim)


def fc2td(fun_fc, f, t, dim):
    """y(t) = 2 Re sum_k F'_k exp(i w_k t), F'_0 = F_0/2 (fc2td.m:24-31)."""
    src = fun_fc if torch.is_tensor(fun_fc) else torch.as_tensor(np.asarray(fun_fc))
    f = _dev(f, torch.float64).reshape(-1)
    t = _dev(t, torch.float64).reshape(-1)
    if dim >= src.dim():
        raise ValueError("Function dimensions and given dimension mismatch")
    if f.numel() != src.shape[dim]:
        raise ValueError("Size along given dimension does not match given fourier coefficients")
    x, shp = _front(_dev(src, torch.complex128), dim)
    ncol = int(np.prod(shp[1:], dtype=np.int64))
    out = torch.empty((t.numel(),) + tuple(shp[1:]), dtype=torch.float64, device=DEV)
    _lib.call("dogpup_fc2td", _lib.ptr(torch.view_as_real(x)), _lib.ptr(out), _lib.ptr(f), _lib.ptr(t),
              f.numel(), t.numel(), ncol, _lib.stream_ptr())
    return torch.movedim(out, 0, dim)


def fc2td_field(field, f, t):
    """fc2td over a whole solver-layout field (fc2td.m:24-31 along the
    frequency axis of phi) on FP64 MFMA: returns float32 [NT][groups][NN][NSP]
    (the time-domain field in the solver's column layout;
    td_field_to_reference gives the reference's NN x NT x NCOLS)."""
    f = _dev(f, torch.float64).reshape(-1)
    t = _dev(t, torch.float64).reshape(-1)
    if f.numel() != field.nf:
        raise ValueError("Size along given dimension does not match given fourier coefficients")
    gpf = field.data.shape[0] // field.nf                   # groups per frequency
    ncol = gpf * field.data.shape[1] * field.data.shape[2]
    out = torch.empty((t.numel(), gpf, field.data.shape[1], field.data.shape[2]), dtype=torch.float32,
                      device=DEV)
    _lib.call("dogpup_fc2td_field", _lib.ptr(torch.view_as_real(field.data)), _lib.ptr(out), _lib.ptr(f),
              _lib.ptr(t), f.numel(), t.numel(), ncol, _lib.stream_ptr())
    return out


def td_field_to_reference(td, ncols):
    """[NT][groups][NN][NSP] -> NN x NT x NCOLS (fc2td(phi, f, t, 2) layout)."""
    nt, gpf, nn, nsp = td.shape
    return td.permute(2, 0, 1, 3).reshape(nn, nt, gpf * nsp)[:, :, :ncols]


def fc2tg(fun_fc, gate_fc, f, nt, dim):
    """y_g = Nt 2 Re sum_k conj(G_gk) F'_k, F'_0 = F_0/2 (fc2tg.m:25-32)."""
    src = fun_fc if torch.is_tensor(fun_fc) else torch.as_tensor(np.asarray(fun_fc))
    f = np.asarray(f if not torch.is_tensor(f) else f.cpu()).reshape(-1)
    if dim >= src.dim():
        raise ValueError("Function dimensions and given dimension mismatch")
    if f.size != src.shape[dim]:
        raise ValueError("Size along given dimension does not match given fourier coefficients")
    g = _dev(gate_fc, torch.complex128).contiguous()
    if g.dim() != 2 or g.shape[1] != f.size:
        raise ValueError("gateFC must be NG x NF")
    x, shp = _front(_dev(src, torch.complex128), dim)
    ncol = int(np.prod(shp[1:], dtype=np.int64))
    out = torch.empty((g.shape[0],) + tuple(shp[1:]), dtype=torch.float64, device=DEV)
    _lib.call("dogpup_fc2tg", _lib.ptr(torch.view_as_real(x)), _lib.ptr(out),
              _lib.ptr(torch.view_as_real(g)), f.size, g.shape[0], ncol, float(nt),
              _lib.stream_ptr())
    return torch.movedim(out, 0, dim)


def id_thresh(data, thresh):
    """Indices of the falling edge at or below ``thresh`` and above 1% of the
    row range (id_thresh.m:13-30); ``'all'`` keeps everything above 1%.
    Rows with zero range normalise to 0 (MATLAB would give NaN)."""
    d = data if torch.is_tensor(data) else torch.as_tensor(np.asarray(data))
    d = d.to(torch.float64)
    lo = d.amin(dim=1, keepdim=True)
    hi = d.amax(dim=1, keepdim=True)
    rng = torch.where(hi - lo == 0, torch.ones_like(hi), hi - lo)
    nd = (d - lo) / rng
    if isinstance(thresh, str):
        if thresh == "all":
            return nd > 0.01
        raise ValueError("thresh must be number or 'all'")
    id_max = torch.argmax(nd, dim=1, keepdim=True)
    cols = torch.arange(d.shape[1], device=d.device)[None, :]
    return (cols >= id_max) & (nd <= thresh) & (nd > 0.01)
