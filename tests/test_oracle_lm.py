"""Pins of the oracle's LM restatement (oracle/lm.py) by what the
mathematics fixes: the spot shape of target_spots (dMesh.m:250-282) and the
LM loop (wghtd_LM_recon.m:23-168) on a closed-form forward model whose
Rytov Jacobian is exact, so the iterates must reach the true absorption."""
import math

import numpy as np
import pytest


def test_target_spots_shape():
    from oracle import lm as olm
    # a line of grid points along x: the Gaussian spot's half-maximum width is
    # (2 sqrt(2 ln 2))^2 w = 8 ln 2 w, because the reference puts the FWHM
    # factor into the standard deviation itself (dMesh.m:264)
    x = np.linspace(-20.0, 20.0, 40001)
    grid = np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1)
    w = 1.5
    t = olm.target_spots(grid, [0.0, 0.0, 0.0], w, "gauss")[0]
    assert t.shape == (x.size,) and abs(t.sum() + 1.0) < 1e-12 and (t <= 0).all()
    above = x[-t >= 0.5 * (-t).max()]
    assert abs((above.max() - above.min()) - 8.0 * math.log(2.0) * w) < 2e-3
    # square spot: the box |x - pos| <= w/2 on every axis, each point -1/count
    g = np.stack(np.meshgrid(np.arange(5.0), np.arange(5.0), np.arange(5.0), indexing="ij"), -1).reshape(-1, 3)
    sq = olm.target_spots(g, [[2.0, 2.0, 2.0], [0.0, 0.0, 0.0]], 2.0, "square")
    assert np.count_nonzero(sq[0]) == 27 and np.allclose(sq[0][sq[0] != 0], -1.0 / 27)
    assert np.count_nonzero(sq[1]) == 8
    inm = np.ones(len(g), bool)
    inm[:10] = False
    assert (olm.target_spots(g, [[2.0, 2.0, 2.0]], 2.0, "square", inm)[0][:10] == 0).all()
    with pytest.raises(ValueError):
        olm.target_spots(g, [[0.0, 0.0, 0.0]], 1.0, "disc")


def _loglinear(seed=0, nm=6, ng=5, nv=4):
    """data(mua) = exp(c + B mua) over (pair, gate): falling curves in the
    gate index (so id_thresh keeps a falling edge), Rytov Jacobian
    d data / d mua / data = B exactly."""
    rng = np.random.default_rng(seed)
    c = np.log(np.linspace(1.0, 0.05, ng))[None, :] + 0.1 * rng.standard_normal((nm, 1))
    B = -300.0 * rng.uniform(0.5, 1.5, (nm, ng, nv))

    def forward(mua):
        d = np.exp(c + np.tensordot(B, mua, axes=(2, 0)))
        return d, B.copy()
    return forward


def test_lm_converges_on_a_loglinear_model():
    """With the exact Jacobian of a log-linear model the weighted LM steps
    (G-normalised update, lambda floored at the misfit) drive mua to the
    truth from a homogeneous start; a wrong sign or a missing G rescaling
    would not."""
    from oracle import lm as olm
    fwd = _loglinear()
    truth = np.array([0.012, 0.008, 0.015, 0.010])
    data_m, _ = fwd(truth)
    mua0 = np.full(4, 0.01)
    r = olm.wghtd_lm_recon(fwd, mua0, data_m, None, (1e-3, 1e-6), max_iter=10)
    # damped by lambda >= |misfit|, every step lowers the misfit and moves
    # toward the truth
    assert r["branch"] and set(r["branch"]) == {"accept"}
    assert all(b < a for a, b in zip(r["errPerc"], r["errPerc"][1:]))
    assert np.linalg.norm(r["mua"] - truth) < 0.7 * np.linalg.norm(mua0 - truth), r["mua"]
    # from a start 2 % off it converges (misfit under 1.5 % of |log data|)
    # onto the truth
    r2 = olm.wghtd_lm_recon(fwd, truth * 1.02, data_m, None, (1e-3, 1e-6), max_iter=10)
    assert r2["exit"] == "converged" and r2["errPerc"][-1] < 1.5
    assert np.abs(r2["mua"] - truth).max() < 5e-3 * truth.max(), r2["mua"]


def test_lm_step_back_branch():
    """A misfit that rises twice in a row takes the step-back branch twice:
    lambda is raised by 10^(1/4) each time (then floored at the misfit) and
    both step-back updates are computed from the cached J and y of the last
    accepted step, so the two are the same vector up to the lambda change."""
    from oracle import lm as olm
    fwd = _loglinear(seed=1)
    truth = np.array([0.012, 0.008, 0.015, 0.010])
    data_m, _ = fwd(truth)
    calls = [0]

    def scripted(mua):
        d, J = fwd(mua)
        s = {2: 3.0, 3: 6.0}.get(calls[0], 1.0)
        calls[0] += 1
        return d * s, J
    r = olm.wghtd_lm_recon(scripted, np.full(4, 0.01), data_m, None, (1e-3, 1e-6), max_iter=5)
    assert r["branch"][:4] == ["accept", "accept", "step back", "step back"]
    k = 10 ** 0.25
    # lambda before the floor: accept divides, step back multiplies
    assert r["lam"][2] >= r["lam"][1] * k * (1 - 1e-12) or r["lam"][2] == pytest.approx(r["err"][2])
    # the two step-back updates: mua_after - mua_before
    h = [np.full(4, 0.01)] + r["mua_hist"]
    u2, u3 = h[3] - h[2], h[4] - h[3]
    cos = u2 @ u3 / (np.linalg.norm(u2) * np.linalg.norm(u3))
    assert cos > 0.99
