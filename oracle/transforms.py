"""Oracle: Fourier-series transforms and data thresholding (test infrastructure only).

Restates ``DOGPUP/td2fc.m``, ``fc2td.m``, ``fc2tg.m`` and ``id_thresh.m``.
The transformed axis is moved to the front, contracted, and moved back, as the
reference does with permute/pagemtimes.
"""
from __future__ import annotations

import numpy as np


def _front(x, dim):
    return np.moveaxis(x, dim, 0)


def _back(x, dim):
    return np.moveaxis(x, 0, dim)


def td2fc(fun_td, f, t, dim):
    """F_k = dt/P * sum_t g(t) exp(-i w_k t), P = t(end) (td2fc.m:26-32)."""
    t = np.asarray(t, float)
    f = np.asarray(f, float)
    if t.size != fun_td.shape[dim]:
        raise ValueError("Size along given dimension does not match given time axis")
    x = _front(np.asarray(fun_td), dim)
    dt = t[1] - t[0]
    P = t[-1]
    T = np.exp(-1j * np.outer(f, t))
    out = dt * np.tensordot(T, x, axes=(1, 0)) / P
    return _back(out, dim)


def fc2td(fun_fc, f, t, dim):
    """y(t) = 2 Re sum_k F'_k exp(i w_k t), F'_0 = F_0/2 (fc2td.m:24-31)."""
    f = np.asarray(f, float)
    t = np.asarray(t, float)
    if f.size != fun_fc.shape[dim]:
        raise ValueError("Size along given dimension does not match given fourier coefficients")
    x = _front(np.array(fun_fc, dtype=complex), dim)
    x[0] = x[0] / 2
    T = np.exp(1j * np.outer(t, f))
    out = 2.0 * np.real(np.tensordot(T, x, axes=(1, 0)))
    return _back(out, dim)


def fc2tg(fun_fc, gate_fc, f, nt, dim):
    """y_g = Nt * 2 Re sum_k conj(G_gk) F'_k, F'_0 = F_0/2 (fc2tg.m:25-32)."""
    f = np.asarray(f, float)
    if f.size != fun_fc.shape[dim]:
        raise ValueError("Size along given dimension does not match given fourier coefficients")
    x = _front(np.array(fun_fc, dtype=complex), dim)
    x[0] = x[0] / 2
    out = nt * 2.0 * np.real(np.tensordot(np.conj(gate_fc), x, axes=(1, 0)))
    return _back(out, dim)


def id_thresh(data, thresh):
    """id_thresh.m:13-30 (rows normalised to [0,1] by their range)."""
    data = np.asarray(data, float)
    lo = data.min(axis=1, keepdims=True)
    hi = data.max(axis=1, keepdims=True)
    rng = np.where(hi - lo == 0, 1.0, hi - lo)
    nd = (data - lo) / rng
    if isinstance(thresh, str):
        if thresh == "all":
            return nd > 0.01
        raise ValueError("thresh must be number or 'all'")
    id_max = np.argmax(nd, axis=1)[:, None]
    cols = np.arange(data.shape[1])[None, :]
    start = (cols >= id_max) & (nd <= thresh)
    end = (cols >= id_max) & (nd > 0.01)
    return start & end


def freq_axis(t_axis, nf):
    """fAxis = (0:Nf-1)*2*pi/T with T = tAxis(end); df = fAxis(2)-fAxis(1)
    (dOptode.m:55-60)."""
    T = float(t_axis[-1])
    f = np.arange(nf) * 2.0 * np.pi / T
    return f, f[1] - f[0]


def source_spectrum(tpsf, t_axis, f_axis, avg_pow):
    """s_tpsf scaled to mean avg_pow, s_fpsf = td2fc(s_tpsf) (dOptode.m:67-73)."""
    tp = np.asarray(tpsf, float)
    tp = tp * (avg_pow / tp.mean())
    return tp, td2fc(tp[None, :], f_axis, t_axis, 1)[0]
