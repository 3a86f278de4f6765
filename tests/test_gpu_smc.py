"""GPU parity of the SMC samplers.

* plain sampler vs the CPU oracle (the same algorithm restated): the
  per-particle arithmetic is identical by construction (shared float
  polynomial math, Philox streams), so log-likelihoods agree to the rounding
  of the FP64 reductions and the resampling decisions agree exactly;
* batched sampler vs the plain sampler: bit-identical log-likelihoods (both
  reduce and scan through the same per-tile functions);
* linear-Gaussian configuration (N = 2^20, T = 100) vs the Kalman filter
  within Monte Carlo error.
"""
import math
import os

import numpy as np
import pytest

import kalman
import oracle as orc

pytestmark = pytest.mark.gpu

MULT, STRAT, SYS = orc.MULTINOMIAL, orc.STRATIFIED, orc.SYSTEMATIC
RES, RSTRAT, RSYS = (orc.RESIDUAL, orc.RESIDUAL_STRATIFIED,
                     orc.RESIDUAL_SYSTEMATIC)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _m():
    import mckl_amd as m
    return m


def _obs():
    return np.load(os.path.join(GOLDEN, "observations.npz"))


CASES = [
    # model, theta, N, seed, scheme
    (1, (0.9, 1.0, 0.5), 4096, 5, SYS),
    (1, (0.9, 1.0, 0.5), 5000, 6, STRAT),
    (1, (0.9, 1.0, 0.5), 3, 7, SYS),
    (1, (0.9, 1.0, 0.5), 2049, 8, SYS),
    (2, (0.95, 0.2, 0.7), 4096, 9, STRAT),
    (2, (0.95, 0.2, 0.7), 70001, 10, SYS),
    # multinomial: sorted uniforms from the log-spacing prefix
    (1, (0.9, 1.0, 0.5), 4096, 11, MULT),
    (2, (0.95, 0.2, 0.7), 5000, 12, MULT),
    (1, (0.9, 1.0, 0.5), 2049, 13, MULT),
    # residual multinomial / stratified / systematic (plain sampler only)
    (1, (0.9, 1.0, 0.5), 4096, 18, RES),
    (2, (0.95, 0.2, 0.7), 70001, 19, RES),
    (1, (0.9, 1.0, 0.5), 2, 20, RES),
    (1, (0.9, 1.0, 0.5), 4096, 14, RSTRAT),
    (2, (0.95, 0.2, 0.7), 70001, 15, RSTRAT),
    (1, (0.9, 1.0, 0.5), 5000, 16, RSYS),
    (2, (0.95, 0.2, 0.7), 3, 17, RSYS),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_plain_vs_oracle(dev, case):
    m = _m()
    model, th, N, seed, scheme = CASES[case]
    y = _obs()["lg_y"][:60] if model == 1 else _obs()["sv_y"][:60]
    s = m.SMCSampler(model, th, y, N, seed, 0.5, scheme)
    s.iterate()
    ll, ess, res = s.results()
    oll, oess, ores = orc.smc_run(model, th, y, N, seed, 0.5, scheme,
                                  history=True)
    print("N=%d ll=%.12f oracle=%.12f resamples=%d" % (N, ll, oll, res.sum()))
    assert np.array_equal(res, ores)
    assert np.allclose(ess, oess, rtol=1e-12, atol=0)
    assert abs(ll - oll) < 1e-10 * max(1.0, abs(oll))


def test_plain_repeatable(dev):
    m = _m()
    y = _obs()["sv_y"][:50]
    a = m.smc_loglik(2, (0.95, 0.2, 0.7), y, 30000, 4, 0.5, STRAT)
    b = m.smc_loglik(2, (0.95, 0.2, 0.7), y, 30000, 4, 0.5, STRAT)
    assert a == b


def test_plain_graph_replays(dev):
    """The sampler replays one captured HIP graph; every replay, with eager
    device work and synchronisations in between, gives the first result."""
    import torch
    m = _m()
    y = _obs()["lg_y"][:80]
    s = m.SMCSampler(1, (0.9, 1.0, 0.5), y, 50000, 21, 0.5, SYS)
    s.iterate()
    first = s.results()
    for k in range(3):
        junk = torch.randn(1 << 20, device=dev).sum()
        torch.cuda.synchronize()
        s.iterate()
        ll, ess, res = s.results()
        assert ll == first[0]
        assert np.array_equal(ess, first[1]) and np.array_equal(res, first[2])
        del junk


def test_lg_config_vs_kalman(dev):
    """Config: N = 2^20 FP32 particles, T = 100 observations simulated from
    seed 7, systematic resampling when ESS < N/2.  Over 64 seeds the mean
    log Z-hat is within 4 standard errors of the Kalman FP64 value (the
    standard error is about 1e-3), and every seed within 6 of the seeds'
    standard deviations of it."""
    m = _m()
    y = _obs()["lg_y"]
    exact = kalman.loglik(y, 0.9, 1.0, 0.5)
    ests = []
    for seed in range(1, 65):
        s = m.SMCSampler(1, (0.9, 1.0, 0.5), y, 1 << 20, seed, 0.5, SYS)
        s.iterate()
        ll, ess, res = s.results()
        ests.append(ll)
        assert 0 < res.sum() < 100
        del s
    ests = np.array(ests)
    sd = ests.std(ddof=1)
    se = sd / np.sqrt(len(ests))
    print("kalman %.6f  smc mean %.6f sd %.5f se %.5f z %.2f" % (
        exact, ests.mean(), sd, se, (ests.mean() - exact) / se))
    assert se < 3e-3
    assert abs(ests.mean() - exact) < 4 * se
    assert np.all(np.abs(ests - exact) < 6 * sd)


def test_batched_equals_plain(dev):
    m = _m()
    ob = _obs()
    theta = ob["batch64_theta"][:24]
    N = ob["batch64_N"][:24].copy()
    N[0] = 3
    N[1] = 2048
    N[2] = 2049
    y = ob["batch64_y"][:24]
    seeds = np.arange(1, 25, dtype=np.uint64)
    b = m.SMCBatch(2, N, theta, y, seeds, 0.5, STRAT)
    for block in (256, 512, 1024):
        b.set_tune(block, 0)
        b.run(use_graph=False)
        lb = b.log_likelihood()
        for f in range(len(N)):
            lp = m.smc_loglik(2, theta[f], y[f], int(N[f]), int(seeds[f]), 0.5, STRAT)
            assert lb[f] == lp, (block, f, lb[f], lp)
    b.run(use_graph=True)
    b.run(use_graph=True)
    assert np.array_equal(b.log_likelihood(), lb)


@pytest.mark.parametrize("scheme", [RES, RSTRAT, RSYS])
def test_residual_batched_equals_plain(dev, scheme):
    """Residual schemes in the batched sampler: bit for bit the plain
    sampler's log-likelihood, for one and two groups per filter (the
    two-group launch has a padding round when a filter's tile count is
    odd) and for filters of 3, 2048, 2049, 6145 and eight more of the
    configuration's sizes (up to 36425 particles)."""
    m = _m()
    ob = _obs()
    theta = ob["batch64_theta"][:12]
    N = ob["batch64_N"][:12].copy()
    N[0] = 3
    N[1] = 2048
    N[2] = 2049
    N[3] = 6145
    y = ob["batch64_y"][:12]
    seeds = np.arange(60, 72, dtype=np.uint64)
    b = m.SMCBatch(2, N, theta, y, seeds, 0.5, scheme)
    lp = np.array([m.smc_loglik(2, theta[f], y[f], int(N[f]), int(seeds[f]),
                                0.5, scheme) for f in range(len(N))])
    for block in (256, 512):
        b.set_tune(block, 0)
        b.run(use_graph=False)
        lb = b.log_likelihood()
        assert np.array_equal(lb, lp), (block, lb - lp)


def test_multinomial_batched_and_kalman(dev):
    """Multinomial resampling: the batched sampler equals the plain one bit
    for bit, and the LG estimate is within Monte Carlo error of Kalman."""
    m = _m()
    ob = _obs()
    theta = ob["batch64_theta"][:8]
    N = ob["batch64_N"][:8]
    y = ob["batch64_y"][:8]
    seeds = np.arange(40, 48, dtype=np.uint64)
    b = m.SMCBatch(2, N, theta, y, seeds, 0.5, MULT)
    b.run()
    lb = b.log_likelihood()
    for f in range(8):
        lp = m.smc_loglik(2, theta[f], y[f], int(N[f]), int(seeds[f]), 0.5, MULT)
        assert lb[f] == lp, (f, lb[f], lp)
    yl = ob["lg_y"]
    exact = kalman.loglik(yl, 0.9, 1.0, 0.5)
    ll = m.smc_loglik(1, (0.9, 1.0, 0.5), yl, 1 << 18, 3, 0.5, MULT)
    print("kalman %.6f multinomial %.6f" % (exact, ll))
    assert abs(ll - exact) < 0.25


@pytest.mark.parametrize("scheme", [RES, RSTRAT, RSYS])
def test_residual_kalman(dev, scheme):
    """Each residual scheme on the LG configuration at N = 2^18: within
    Monte Carlo error of the Kalman log-likelihood, and every resampling
    step keeps the integral parts (checked through the oracle run above at
    small sizes; here through the estimate)."""
    m = _m()
    y = _obs()["lg_y"]
    exact = kalman.loglik(y, 0.9, 1.0, 0.5)
    ll = m.smc_loglik(1, (0.9, 1.0, 0.5), y, 1 << 18, 5, 0.5, scheme)
    print("kalman %.6f scheme %d %.6f" % (exact, scheme, ll))
    assert abs(ll - exact) < 0.25


def test_batched_vs_oracle(dev):
    m = _m()
    ob = _obs()
    theta = ob["batch64_theta"][:6]
    N = ob["batch64_N"][:6]
    y = ob["batch64_y"][:6]
    seeds = np.arange(11, 17, dtype=np.uint64)
    b = m.SMCBatch(2, N, theta, y, seeds, 0.5, STRAT)
    b.run()
    lb = b.log_likelihood()
    for f in range(6):
        lo = orc.smc_run(2, theta[f], y[f], int(N[f]), int(seeds[f]), 0.5, STRAT)
        assert abs(lb[f] - lo) < 1e-10 * max(1.0, abs(lo))


def test_smc_errors(dev):
    m = _m()
    with pytest.raises(m.MCKLError):
        m.SMCSampler(1, (0.9, 1.0, 0.5), np.zeros(3), 0, 1)
    with pytest.raises(m.MCKLError):
        m.SMCSampler(2, (1.5, 0.2, 0.7), np.zeros(3), 100, 1)
    with pytest.raises(m.MCKLError):
        m.SMCSampler(1, (0.9, 1.0, 0.5), np.zeros(3), 100, 1, 0.5, 7)
    with pytest.raises(m.MCKLError):
        m.SMCBatch(1, [100, 200], [(0.9, 1.0, 0.5)] * 2, np.zeros((2, 3)),
                   [1, 2], 0.5, 6)


def test_residual_systematic_equals_systematic(dev):
    """Residual systematic resampling gives the replication numbers of
    systematic resampling with the same uniform (sum_k<=i I_k is an
    integer, so floor(N P_i - u) splits into the integral parts and the
    residual count): the samplers agree run for run, at any size."""
    m = _m()
    y = _obs()["lg_y"][:40]
    for N in (3, 4096, 20011):
        a = m.SMCSampler(1, (0.9, 1.0, 0.5), y, N, 21, 0.5, SYS)
        b = m.SMCSampler(1, (0.9, 1.0, 0.5), y, N, 21, 0.5, RSYS)
        a.iterate()
        b.iterate()
        la, ea, ra = a.results()
        lb, eb, rb = b.results()
        assert np.array_equal(ra, rb)
        assert abs(la - lb) < 1e-10 * abs(la), (N, la, lb)


@pytest.mark.parametrize("N,T,scheme", [((1 << 21) + 77, 8, SYS),
                                        (1, 12, STRAT), (2, 12, MULT),
                                        ((1 << 21) + 77, 4, RES),
                                        (2048, 1, SYS),
                                        ((1 << 23) + 5, 3, STRAT),
                                        ((1 << 24) + 5, 3, SYS)])
def test_plain_sizes_vs_oracle(dev, N, T, scheme):
    """Sizes the configurations do not reach: more than 1024 tiles (several
    tiles per thread in every cross-tile reduction; at 2^24 + 5, 8193
    tiles: 33 per thread, a look-back of nine rounds and more workgroups
    than the GPU holds at once), one and two particles, a single step."""
    m = _m()
    y = _obs()["lg_y"][:T]
    th = (0.9, 1.0, 0.5)
    # threshold 0.99: every step with unequal weights resamples
    s = m.SMCSampler(1, th, y, N, 31, 0.99, scheme)
    s.iterate()
    ll, ess, res = s.results()
    oll, oess, ores = orc.smc_run(1, th, y, N, 31, 0.99, scheme, nthreads=16,
                                  history=True)
    if N > 2:
        assert res.sum() == T
    print("N=%d ll=%.12f oracle=%.12f resamples=%d" % (N, ll, oll, res.sum()))
    assert np.array_equal(res, ores)
    assert np.allclose(ess, oess, rtol=1e-11, atol=0)
    assert abs(ll - oll) < 1e-9 * max(1.0, abs(oll))


def test_batched_extreme_sizes(dev):
    """Filters of 1, 2 and the maximum 131072 particles in one batch equal
    their plain runs bit for bit."""
    m = _m()
    ob = _obs()
    N = np.array([1, 2, 131072, 2047, 2048], np.uint32)
    theta = ob["batch64_theta"][:5]
    y = ob["batch64_y"][:5, :20]
    seeds = np.arange(90, 95, dtype=np.uint64)
    for scheme in (MULT, STRAT, SYS):
        b = m.SMCBatch(2, N, theta, y, seeds, 0.5, scheme)
        b.run()
        lb = b.log_likelihood()
        for f in range(5):
            lp = m.smc_loglik(2, theta[f], y[f], int(N[f]), int(seeds[f]), 0.5,
                              scheme)
            assert lb[f] == lp, (scheme, f, lb[f], lp)


def test_results_wait_for_their_own_run(dev):
    """A run enqueued on a side stream, read right away: the result call
    waits for that run's completion event (not the whole device) and
    returns the same bits as a run on the default stream; so does the
    batched form."""
    import torch
    m = _m()
    y = np.linspace(-1.0, 1.0, 40)
    want = m.smc_loglik(m.MODEL_LG, (0.9, 1.0, 0.5), y, 50000, seed=3)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        s = m.SMCSampler(m.MODEL_LG, (0.9, 1.0, 0.5), y, 50000, seed=3)
        s.iterate()
        got = s.log_likelihood()
    assert got == want
    N = np.array([3000, 700], np.uint32)
    th = np.array([[0.95, 0.2, 0.7], [0.9, 0.25, 0.6]])
    yy = np.tile(y[:20], (2, 1))
    b0 = m.SMCBatch(m.MODEL_SV, N, th, yy, [5, 6])
    b0.run()
    with torch.cuda.stream(side):
        b1 = m.SMCBatch(m.MODEL_SV, N, th, yy, [5, 6])
        b1.run()
        r1 = b1.log_likelihood()
    assert np.array_equal(r1, b0.log_likelihood())


TUNE_CANDIDATES = [(256, 0, 0, 256), (256, 2, 0, 256), (512, 0, 0, 256),
                   (512, 2, 0, 256), (1024, 0, 0, 256), (512, 0, 2, 256),
                   (512, 0, 4, 256), (512, 0, 8, 256), (1024, 0, 4, 256),
                   (1024, 0, 8, 256), (1024, 0, 16, 256), (1024, 0, 24, 256),
                   (512, 2, 8, 256), (1024, 2, 8, 256), (256, 0, 2, 512),
                   (1024, 0, 1, 1024)]


def test_every_tuned_configuration_equals_plain(dev):
    """Every configuration the autotuner tries (and a few more: a wide small
    launch, every filter in the large launch) gives each filter the plain
    sampler's bits, graph or not: 40 SV filters of 1 to 40 tiles."""
    m = _m()
    rng = np.random.default_rng(17)
    nf = 40
    N = np.concatenate([[3, 2048, 2049, 40 * 2048],
                        rng.integers(256, 40 * 2048, nf - 4)]).astype(np.uint32)
    theta = np.stack([rng.uniform(0.9, 0.99, nf), rng.uniform(0.1, 0.3, nf),
                      rng.uniform(0.5, 1.0, nf)], 1)
    y = rng.normal(size=(nf, 12)) * 0.8
    seeds = np.arange(100, 100 + nf, dtype=np.uint64)
    plain = np.array([m.smc_loglik(2, theta[f], y[f], int(N[f]),
                                   int(seeds[f]), 0.5, STRAT)
                      for f in range(nf)])
    b = m.SMCBatch(2, N, theta, y, seeds, 0.5, STRAT)
    for cand in TUNE_CANDIDATES:
        b.set_tune(*cand)
        for g in (False, True):
            b.run(use_graph=g)
            assert np.array_equal(b.log_likelihood(), plain), (cand, g)


def test_autotune_and_cache(dev, tmp_path):
    """The library's autotuner picks one of its configurations, checks the
    bits of every one, records the choice, and a second batch of the same
    workload takes it from the cache without timing runs."""
    m = _m()
    ob = _obs()
    theta = ob["batch64_theta"]
    N = ob["batch64_N"]
    y = ob["batch64_y"]
    seeds = np.arange(1, len(N) + 1, dtype=np.uint64)
    cache = str(tmp_path / "tune.jsonl")
    b = m.SMCBatch(2, N, theta, y, seeds, 0.5, STRAT)
    tune, ms = b.autotune(cache)
    assert tuple(tune.values()) in [tuple(c) for c in TUNE_CANDIDATES]
    assert ms > 0
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and '"block": %d' % tune["block"] in lines[0]
    b.run()
    lb = b.log_likelihood()
    b2 = m.SMCBatch(2, N, theta, y, seeds, 0.5, STRAT)
    tune2, ms2 = b2.autotune(cache)
    assert tune2 == tune and ms2 == pytest.approx(ms, rel=1e-3)
    assert len(open(cache).read().splitlines()) == 1
    b2.run()
    assert np.array_equal(b2.log_likelihood(), lb)
    # another workload (other sizes) is not served by that record
    b3 = m.SMCBatch(2, N[:10], theta[:10], y[:10], seeds[:10], 0.5, STRAT)
    b3.autotune(cache)
    assert len(open(cache).read().splitlines()) == 2


def test_batched_config4_equals_plain_runs(dev):
    """The batched configuration at full size -- 4096 stochastic-volatility
    filters, N log-uniform in [256, 65536], T = 50, stratified -- as one
    graph launch with the autotuned configuration, against one plain
    sampler run per filter: all 4096 log-likelihoods bit-identical."""
    m = _m()
    from mckl_amd import workloads as wl
    theta, Nf, yb, seeds = wl.sv_batch(4096, 50)
    b = m.SMCBatch(m.MODEL_SV, Nf, theta, yb, seeds, 0.5, STRAT)
    tune, ms = b.autotune()
    print("autotuned %s: %.3f ms per run" % (tune, ms))
    b.run(use_graph=True)
    lb = b.log_likelihood()
    lp = np.empty(len(Nf))
    for f in range(len(Nf)):
        s = m.SMCSampler(m.MODEL_SV, theta[f], yb[f], int(Nf[f]),
                         int(seeds[f]), 0.5, STRAT)
        s.iterate()
        lp[f] = s.log_likelihood()
    assert np.array_equal(lb, lp), int((lb != lp).sum())
