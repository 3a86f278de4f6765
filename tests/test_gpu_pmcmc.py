"""PMMH on the device log-likelihood (algorithm/pmcmc.hpp:71-158).

Chains of the single mutation and of the lockstep batch against the
oracle's restatement driven by the same engine streams: the acceptance
draws are bit-identical, the random-walk normals agree to the Normal fill's
2 ulp and the log-likelihoods to ~1e-12, so the decisions are the same and
the parameters agree to rounding; and the batch's posterior of the linear-Gaussian
autoregression coefficient against the exact posterior by quadrature of the
Kalman likelihood."""
import math
import os

import numpy as np
import pytest

import kalman
import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _m():
    import mckl_amd as m
    return m


def _y():
    return np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"]


def prior_box(P):
    P = np.atleast_2d(P)
    ok = (np.abs(P[:, 0]) < 1) & (P[:, 1] > 0) & (P[:, 2] > 0)
    return np.where(ok, 0.0, -np.inf)


def test_single_chain_matches_oracle(dev):
    from mckl_amd import pmcmc
    m = _m()
    y = _y()
    mut = pmcmc.PMCMCMutation(256, 20, lambda p: float(prior_box(p)[0]),
                              m.MODEL_LG, y, seed=3)
    mut.update(pmcmc.random_walk((0.05, 0.05, 0.05)))
    param = np.array([0.8, 1.0, 0.5])
    acc = [mut(it, param) for it in range(30)]
    P = np.array([[0.8, 1.0, 0.5]])
    oacc, olnc = orc.pmmh(orc.MODEL_LG, y, 256, 20, prior_box,
                          [orc.pmmh_random_walk((0.05, 0.05, 0.05))], P, 30,
                          seed=3)
    assert acc == oacc[:, 0].tolist()
    assert np.allclose(param, P[0], rtol=1e-13, atol=0)
    assert abs(mut.log_nc - olnc[0]) < 1e-9 * abs(olnc[0])
    assert 0 < sum(acc) < 29


def test_batch_matches_oracle(dev):
    from mckl_amd import pmcmc
    m = _m()
    y = _y()
    C = 5
    bt = pmcmc.PMCMCBatch(200, 15, prior_box, m.MODEL_LG, y, C, seed=11)
    bt.update(pmcmc.random_walk((0.1, 0.1, 0.1)))
    params = np.tile([0.5, 0.8, 0.7], (C, 1))
    acc = np.array([bt(it, params) for it in range(20)])
    P = np.tile([0.5, 0.8, 0.7], (C, 1))
    oacc, olnc = orc.pmmh(orc.MODEL_LG, y, 200, 15, prior_box,
                          [orc.pmmh_random_walk((0.1, 0.1, 0.1))], P, 20,
                          seed=11)
    assert np.array_equal(acc, oacc)
    assert np.allclose(params, P, rtol=1e-13, atol=0)
    fin = np.isfinite(olnc)
    assert np.allclose(bt.log_nc[fin], olnc[fin], rtol=1e-9, atol=0)


def test_batch_posterior_vs_quadrature(dev):
    """Flat prior on a in (-1, 1) with sx = 1, sy = 0.5 fixed; 256 chains,
    120 iterations after 40 of burn-in, N = 512 particles, T = 50."""
    from mckl_amd import pmcmc
    m = _m()
    y = _y()[:50]
    grid = np.linspace(-0.999, 0.999, 4001)
    ll = np.array([kalman.loglik(y, a, 1.0, 0.5) for a in grid])
    w = np.exp(ll - ll.max())
    w /= w.sum()
    mean = float(w @ grid)
    sd = math.sqrt(float(w @ (grid - mean) ** 2))
    C = 256
    bt = pmcmc.PMCMCBatch(512, 50, prior_box, m.MODEL_LG, y, C, seed=5)
    bt.update(pmcmc.random_walk((0.08, 0.0, 0.0)))
    params = np.tile([0.0, 1.0, 0.5], (C, 1))
    params[:, 0] = np.linspace(-0.5, 0.9, C)
    draws = []
    accs = 0
    for it in range(160):
        a = bt(it, params)
        if it >= 40:
            draws.append(params[:, 0].copy())
            accs += a.sum()
    D = np.array(draws)                      # (iters, chains)
    chain_means = D.mean(0)
    se = chain_means.std(ddof=1) / math.sqrt(C)
    assert abs(D.mean() - mean) < 5 * se + 1e-3, (D.mean(), mean, se)
    assert abs(D.std() - sd) < 0.2 * sd
    assert 0.1 < accs / D.size < 0.9
