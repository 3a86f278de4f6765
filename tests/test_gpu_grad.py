"""GPU checks of the SMC log-likelihood gradient (hand-written adjoint).

* the replayed function at the forward point equals the forward estimate,
  bit for bit (same kernel, same ancestors);
* the adjoint gradient equals central finite differences (h = 1e-4) of the
  replayed function: the ancestors and resampling decisions are held fixed
  and the noise is common (counter-based), so the replayed function is
  smooth and the differences are exact up to O(h^2) and rounding;
* the checkpoint interval K does not change a bit of the result;
* on the linear-Gaussian model the mean gradient over independent seeds
  matches the Kalman FP64 gradient within Monte Carlo error, and within
  1e-3 relative where the Monte Carlo error allows it (T = 10, 16384 seeds,
  standard errors below 3e-4 relative).
"""
import os

import numpy as np
import pytest

import kalman

pytestmark = pytest.mark.gpu

MULT, STRAT, SYS = 0, 1, 2  # ResampleScheme numbering
RES, RES_STRAT, RES_SYS = 3, 4, 5  # their residual forms

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LG = (0.9, 1.0, 0.5)
SV = (0.95, 0.2, 0.7)


def _m():
    import mckl_amd as m
    return m


def _obs():
    return np.load(os.path.join(GOLDEN, "observations.npz"))


def _fd(g, th, h=1e-4):
    """Central differences of replay() along each parameter, all points."""
    th = np.asarray(th, np.float64).reshape(-1, 3)
    out = np.zeros_like(th)
    for j in range(3):
        tp = th.copy()
        tm = th.copy()
        tp[:, j] += h
        tm[:, j] -= h
        out[:, j] = (g.replay(tp) - g.replay(tm)) / (2 * h)
    return out


@pytest.mark.parametrize("model,N,K,scheme", [
    (1, 3000, 7, SYS), (2, 4096, 10, STRAT), (2, 2049, 60, SYS),
    (1, 5, 3, SYS), (1, 3000, 9, MULT), (2, 4100, 13, MULT),
    (1, 3000, 7, RES_SYS), (2, 4100, 11, RES_STRAT), (2, 2049, 9, RES),
    (1, 7, 3, RES)])
def test_replay_and_fd(dev, model, N, K, scheme):
    m = _m()
    y = _obs()["lg_y"][:60] if model == 1 else _obs()["sv_y"][:60]
    th = np.array([LG if model == 1 else SV] * 3)
    th[1, 0] -= 0.02
    th[2, 1] += 0.05
    g = m.SMCGradient(model, 3, N, y, [3, 4, 5], 0.5, scheme, K=K, tape=True)
    ll, gr = g.run(th)
    v = g.replay(th)
    print("N=%d K=%d ll=%s grad=%s" % (N, K, ll, gr.tolist()))
    assert np.array_equal(v, ll)
    fd = _fd(g, th, 1e-4)
    rel = np.abs(fd - gr) / np.maximum(np.abs(gr), 1.0)
    print("fd(h=1e-4)=%s max rel %.3g" % (fd.tolist(), rel.max()))
    assert rel.max() < 1e-3
    # Richardson extrapolation of the central differences: O(h^4) error,
    # which separates truncation error (the states are polynomials of high
    # degree in a) from a missing term in the adjoint
    fd2 = _fd(g, th, 5e-5)
    rich = (4 * fd2 - fd) / 3
    rel = np.abs(rich - gr) / np.maximum(np.abs(gr), 1.0)
    print("richardson=%s max rel %.3g" % (rich.tolist(), rel.max()))
    assert rel.max() < 2e-6


def test_checkpoint_interval_invariance(dev):
    m = _m()
    y = _obs()["sv_y"][:40]
    for scheme in (SYS, MULT, RES_SYS, RES):
        res = []
        for K in (1, 3, 13, 40, 100):
            g = m.SMCGradient(2, 2, 5000, y, [8, 9], 0.5, scheme, K=K)
            res.append(g.run([SV, (0.9, 0.25, 0.6)]))
        for ll, gr in res[1:]:
            assert np.array_equal(ll, res[0][0])
            assert np.array_equal(gr, res[0][1])


def test_per_point_observations_and_one_shot(dev):
    m = _m()
    ob = _obs()
    th = ob["batch64_theta"][:5]
    y = ob["batch64_y"][:5]
    seeds = np.arange(30, 35, dtype=np.uint64)
    g = m.SMCGradient(2, 5, 3000, y, seeds, 0.5, STRAT, K=10)
    ll, gr = g.run(th)
    ll1, gr1 = m.smc_loglik_grad(2, th, y, 3000, seeds, 0.5, STRAT, K=10)
    assert np.array_equal(ll, ll1) and np.array_equal(gr, gr1)
    for p in range(5):
        gp = m.SMCGradient(2, 1, 3000, y[p], [seeds[p]], 0.5, STRAT, K=25)
        llp, grp = gp.run(th[p])
        assert llp[0] == ll[p] and np.array_equal(grp[0], gr[p])


@pytest.mark.parametrize("scheme", [SYS, RES_SYS])
def test_forward_close_to_plain_sampler(dev, scheme):
    """FP64 forward of the gradient sampler vs the FP32 plain sampler on the
    same noise: without resampling (threshold 0) the two log-likelihoods
    agree up to the FP32 rounding of the plain states over 15 steps.  With
    resampling a replication count can flip at a boundary between the two
    precisions and send the runs down different ancestries (a Monte Carlo
    difference), so there the paired differences over 16 seeds are only
    required to centre on zero within four standard errors."""
    m = _m()
    y = _obs()["lg_y"][:15]
    N = 4096
    for seed in (1, 2, 3):
        g = m.SMCGradient(1, 1, N, y, [seed], 0.0, SYS, K=5)
        ll, _ = g.run(LG)
        lp = m.smc_loglik(1, LG, y, N, seed, 0.0, SYS)
        print("seed %d no resampling: fp64 %.9f fp32 %.9f" % (seed, ll[0], lp))
        assert abs(ll[0] - lp) < 1e-4
    seeds = np.arange(1, 17)
    g = m.SMCGradient(1, len(seeds), N, y, seeds, 0.5, scheme, K=5)
    ll, _ = g.run(np.tile(LG, (len(seeds), 1)))
    lp = np.array([m.smc_loglik(1, LG, y, N, int(s), 0.5, scheme)
                   for s in seeds])
    d = ll - lp
    print("paired differences with resampling:", np.round(d, 5))
    assert abs(d.mean()) < 4 * max(d.std(ddof=1), 1e-6) / np.sqrt(len(d))


@pytest.mark.parametrize("scheme", [SYS, RES_STRAT])
def test_lg_mean_gradient_vs_kalman(dev, scheme):
    """Linear-Gaussian (0.9, 1, 0.5), T = 100 observations from seed 7,
    N = 2^16, 64 seeds: |mean - Kalman| within 4 standard errors."""
    m = _m()
    y = _obs()["lg_y"]
    nb = 64
    g = m.SMCGradient(1, nb, 1 << 16, y, np.arange(1, nb + 1), 0.5, scheme,
                      K=50)
    ll, gr = g.run(LG)
    ex = kalman.grad(y, *LG)
    mean = gr.mean(0)
    se = gr.std(0, ddof=1) / np.sqrt(nb)
    z = (mean - ex) / se
    print("kalman %s mean %s se %s z %s rel %s" % (
        ex, mean, se, z, np.abs(mean - ex) / np.abs(ex)))
    assert np.all(np.abs(z) < 4.5)
    assert np.all(np.abs(mean - ex) < 0.01 * np.abs(ex) + 4.5 * se)


def test_lg_gradient_within_1e3_of_kalman(dev):
    """T = 10 of the linear-Gaussian observations at theta = (0.7, 1.3, 0.7)
    (away from the generating point, so the gradient is large against its
    spread), N = 2^16, systematic resampling, 16384 seeds in one batch: the
    mean gradient is within 1e-3 relative of the Kalman FP64 gradient in
    every component, with standard errors below 3e-4 relative so that the
    bound is resolved.  The per-seed spread grows about linearly in T
    (path degeneracy), which is why the series is short."""
    m = _m()
    y = _obs()["lg_y"][:10]
    th = (0.7, 1.3, 0.7)
    nb = 16384
    g = m.SMCGradient(1, nb, 1 << 16, y, np.arange(1, nb + 1), 0.5, SYS, K=5)
    _, gr = g.run(th)
    ex = kalman.grad(y, *th)
    mean = gr.mean(0)
    se = gr.std(0, ddof=1) / np.sqrt(nb)
    rel = np.abs(mean - ex) / np.abs(ex)
    print("kalman %s mean %s rel err %s se rel %s" % (
        ex, mean, rel, se / np.abs(ex)))
    assert np.all(se / np.abs(ex) < 3e-4)
    assert np.all(rel < 1e-3)


def test_grad_errors(dev):
    m = _m()
    with pytest.raises(m.MCKLError):
        m.SMCGradient(1, 1, 1 << 18, np.zeros(3), [1])
    with pytest.raises(m.MCKLError):
        m.SMCGradient(1, 1, 100, np.zeros(3), [1], scheme=6)
    g = m.SMCGradient(2, 1, 100, np.zeros(3), [1])
    with pytest.raises(m.MCKLError):
        g.replay(SV)
    with pytest.raises(m.MCKLError):
        g.run((1.2, 0.2, 0.7))


def test_torch_autograd_wrapper(dev):
    """The adjoint behind a torch.autograd.Function (test harness only):
    backward of sum(w * loglik) gives w[:, None] * grad, and autograd's own
    numerical check of the replayed function agrees with it."""
    import torch
    m = _m()
    y = _obs()["sv_y"][:30]
    g = m.SMCGradient(2, 2, 2000, y, [5, 6], 0.5, STRAT, K=8, tape=True)

    class LogLik(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            ll, gr = g.run(theta.detach().cpu().numpy())
            ctx.save_for_backward(torch.from_numpy(gr))
            return torch.from_numpy(ll)

        @staticmethod
        def backward(ctx, go):
            (gr,) = ctx.saved_tensors
            return go[:, None] * gr

    th = torch.tensor([SV, (0.9, 0.25, 0.6)], dtype=torch.float64,
                      requires_grad=True)
    w = torch.tensor([0.3, -1.7], dtype=torch.float64)
    (w * LogLik.apply(th)).sum().backward()
    _, gr = g.run(th.detach().numpy())
    assert torch.allclose(th.grad, w[:, None] * torch.from_numpy(gr),
                          rtol=0, atol=0)
    # central differences of the replayed function through torch
    h = 1e-5
    for j in range(3):
        e = np.zeros((2, 3))
        e[:, j] = h
        base = th.detach().numpy()
        d = (g.replay(base + e) - g.replay(base - e)) / (2 * h)
        assert np.allclose(d, gr[:, j], rtol=1e-5, atol=1e-6)
