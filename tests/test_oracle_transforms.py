"""Pins the transform oracle (oracle/transforms.py): closed forms and the
fc2tg / fc2td / td2fc identity."""
import numpy as np
import pytest

from conftest import rel
from oracle import transforms as otr


def test_td2fc_gaussian_closed_form():
    t = np.arange(0, 2001) * 10e-12            # P = 20 ns
    f, df = otr.freq_axis(t, 40)
    sd, t0 = 150e-12, 2e-9
    g = np.exp(-((t - t0) ** 2) / (2 * sd ** 2))
    F = otr.td2fc(g[None, :], f, t, 1)[0]
    P = t[-1]
    exact = sd * np.sqrt(2 * np.pi) / P * np.exp(-1j * f * t0) * np.exp(-(f * sd) ** 2 / 2)
    assert rel(F, exact) < 1e-10
    # df = 2 pi / t(end) (dOptode.m:57-60)
    assert df == pytest.approx(2 * np.pi / P)


def test_fc2tg_equals_gated_sum_of_fc2td():
    rng = np.random.default_rng(0)
    nt, nf = 501, 12
    t = np.arange(nt) * 20e-12
    f, _ = otr.freq_axis(t, nf)
    F = rng.normal(size=(5, nf)) + 1j * rng.normal(size=(5, nf))
    gates = rng.uniform(size=(3, nt))
    G = otr.td2fc(gates, f, t, 1)
    tg = otr.fc2tg(F, G, f, nt, 1)
    td = otr.fc2td(F, f, t, 1)                  # 5 x nt
    expect = nt / (nt - 1) * td @ gates.T       # (Nt/(Nt-1)) sum_t g(t) y(t)
    assert rel(tg, expect) < 1e-12


def test_fc2td_reconstructs_band_limited_signal():
    nt = 1001
    t = np.arange(nt) * 20e-12
    f, _ = otr.freq_axis(t, 16)
    P = t[-1]
    g = 1.5 + np.cos(f[3] * t + 0.3) + 0.5 * np.sin(f[7] * t)
    F = otr.td2fc(g[None, :], f, t, 1)
    y = otr.fc2td(F, f, t, 1)[0]
    # td2fc sums both end samples of the period: O(NF dt / P) deviation
    assert np.abs(y - g).max() < 4 * 16 * (t[1] - t[0]) / P * 3


def test_dim_handling_and_dc_halving():
    nt = 11
    t = np.arange(nt) * 1.0
    f, _ = otr.freq_axis(t, 3)
    F = np.zeros((2, 3, 4), complex)
    F[:, 0, :] = 2.0                            # DC only: fc2td gives F0 (halved, doubled)
    y = otr.fc2td(F, f, t, 1)
    assert y.shape == (2, nt, 4) and np.allclose(y, 2.0)


def test_id_thresh_falling_edge():
    d = np.array([[0.0, 0.2, 1.0, 0.9, 0.7, 0.3, 0.005, 0.0]])
    ids = otr.id_thresh(d, 0.8)
    assert ids.tolist() == [[False, False, False, False, True, True, False, False]]
    assert otr.id_thresh(d, "all").tolist() == [[False, True, True, True, True, True, False, False]]
