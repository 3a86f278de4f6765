"""Pin the SMC oracle: the float math functions against float64 closed
forms, the Kalman oracle against the dense Gaussian density, and the SMC
restatement against the Kalman log-likelihood within Monte Carlo error.
CPU only."""
import ctypes
import math
import os

import numpy as np
import pytest

import kalman
import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

O = orc.O
O.mo_log_f.restype = ctypes.c_float
O.mo_log_f.argtypes = [ctypes.c_float]
O.mo_exp_f.restype = ctypes.c_float
O.mo_exp_f.argtypes = [ctypes.c_float]
O.mo_sincos2pi_f.argtypes = [ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                             ctypes.POINTER(ctypes.c_float)]


def _ulp(a, ref):
    a = np.asarray(a, np.float32)
    return np.abs(a.astype(np.float64) - ref) / np.spacing(
        np.abs(ref).astype(np.float32)).astype(np.float64)


def test_float_math_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.random(20000).astype(np.float32),
                         (np.arange(1, 4096) * 2.0 ** -24).astype(np.float32),
                         (rng.random(2000) * 100).astype(np.float32)])
    lf = np.array([O.mo_log_f(float(x)) for x in xs], np.float32)
    ref = np.log(xs.astype(np.float64))
    m = ref != 0
    assert _ulp(lf[m], ref[m]).max() < 1.0
    assert O.mo_log_f(1.0) == 0.0
    es = np.concatenate([-(rng.random(20000) * 87).astype(np.float32),
                         (rng.random(2000) * 10 - 5).astype(np.float32)])
    ef = np.array([O.mo_exp_f(float(x)) for x in es], np.float32)
    assert _ulp(ef, np.exp(es.astype(np.float64))).max() < 1.0
    assert O.mo_exp_f(0.0) == 1.0 and O.mo_exp_f(-90.0) == 0.0
    s, c = ctypes.c_float(), ctypes.c_float()
    for u in list(rng.random(5000).astype(np.float32)) + [0.0, 0.25, 0.5, 1.0]:
        O.mo_sincos2pi_f(float(u), ctypes.byref(s), ctypes.byref(c))
        th = 2 * math.pi * float(u)
        assert abs(s.value - math.sin(th)) < 1e-7
        assert abs(c.value - math.cos(th)) < 1e-7


def test_smc_normals_are_standard():
    z = np.array([orc.smc_normal(9, i, 3) for i in range(200000)])
    n = len(z)
    assert abs(z.mean()) < 5 / math.sqrt(n)
    assert abs(z.var() - 1) < 5 * math.sqrt(2.0 / n)
    # even/odd particles share a block: cos and sin branches are independent
    assert abs(np.corrcoef(z[0::2], z[1::2])[0, 1]) < 5 / math.sqrt(n / 2)


def test_kalman_matches_dense_gaussian():
    rng = np.random.default_rng(1)
    for T in (1, 2, 5, 40):
        y = rng.normal(size=T)
        for th in ((0.9, 1.0, 0.5), (0.5, 0.3, 2.0), (-0.7, 1.5, 0.1)):
            a = kalman.loglik(y, *th)
            b = kalman.loglik_dense(y, *th)
            assert abs(a - b) < 1e-9 * max(1.0, abs(b))


def test_kalman_gradient_matches_finite_differences():
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"]
    th = np.array([0.9, 1.0, 0.5])
    g = kalman.grad(y, *th)
    for k in range(3):
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd = (kalman.loglik(y, *tp) - kalman.loglik(y, *tm)) / (2 * h)
        assert abs(fd - g[k]) < 1e-5 * max(1.0, abs(g[k]))


def test_kalman_fixture_of_the_bench():
    """tests/golden/kalman_lg.json (the bench's reference values) holds the
    Kalman log-likelihood and gradient of the committed observations."""
    import json
    ref = json.load(open(os.path.join(GOLDEN, "kalman_lg.json")))
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"]
    for r in ref.values():
        T = r["T"]
        assert np.array_equal(np.asarray(r["y"]), y[:T])
        assert r["loglik"] == kalman.loglik(y[:T], *r["theta"])
        assert np.array_equal(np.asarray(r["grad"]),
                              kalman.grad(y[:T], *r["theta"]))


def test_smc_oracle_vs_kalman():
    """Bootstrap PF log-likelihood on the LG configuration (T = 100,
    observations from the committed fixture) within Monte Carlo error of the
    exact Kalman value."""
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"]
    exact = kalman.loglik(y, 0.9, 1.0, 0.5)
    ests = [orc.smc_run(orc.MODEL_LG, (0.9, 1.0, 0.5), y, 1 << 14, seed,
                        0.5, orc.SYSTEMATIC, nthreads=4) for seed in range(1, 17)]
    ests = np.array(ests)
    sd = ests.std(ddof=1)
    # log Z-hat has a small negative bias (about var/2); allow 4 standard
    # errors plus that bias.  Its spread at N = 2^14 is about 0.19 (40
    # seeds); 16 seeds estimate it to within about 0.035.
    assert abs(ests.mean() - exact) < 4 * sd / math.sqrt(len(ests)) + sd * sd
    assert sd < 0.3


@pytest.mark.parametrize("scheme", [orc.SYSTEMATIC, orc.STRATIFIED,
                                    orc.MULTINOMIAL, orc.RESIDUAL,
                                    orc.RESIDUAL_STRATIFIED,
                                    orc.RESIDUAL_SYSTEMATIC])
def test_smc_oracle_schemes_consistent(scheme):
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"][:40]
    exact = kalman.loglik(y, 0.9, 1.0, 0.5)
    ll, ess, res = orc.smc_run(orc.MODEL_LG, (0.9, 1.0, 0.5), y, 1 << 13, 3,
                               0.5, scheme, nthreads=4, history=True)
    assert abs(ll - exact) < 0.3
    assert res.sum() > 0 and np.all(ess <= (1 << 13) + 1e-6)


def test_kalman_filter_moments_match_dense_gaussian():
    rng = np.random.default_rng(2)
    for T in (1, 3, 25):
        y = rng.normal(size=T)
        for th in ((0.9, 1.0, 0.5), (-0.6, 0.4, 1.5)):
            m, P = kalman.filter_moments(y, *th)
            md, Pd = kalman.filter_moments_dense(y, *th)
            assert np.allclose(m, md, rtol=1e-9, atol=1e-10)
            assert np.allclose(P, Pd, rtol=1e-9, atol=1e-12)


def test_smc_oracle_moments_vs_kalman():
    """Selection estimates of E[x_t | y_1..t] and E[x_t^2 | y_1..t] from the
    oracle sampler within Monte Carlo error of the Kalman filter."""
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"][:40]
    th = (0.9, 1.0, 0.5)
    m, P = kalman.filter_moments(y, *th)
    N = 1 << 15
    for seed in (1, 2):
        ll, ess, res, est = orc.smc_run(orc.MODEL_LG, th, y, N, seed, 0.5,
                                        orc.SYSTEMATIC, nthreads=4, moments=2)
        assert est.shape == (40, 2)
        se = np.sqrt(P / ess)
        assert np.all(np.abs(est[:, 0] - m) < 8 * se + 1e-3)
        assert np.all(np.abs(est[:, 1] - (P + m * m)) <
                      8 * se * (2 * np.abs(m) + 2 * np.sqrt(P)) + 1e-3)
        # the estimator does not disturb the run
        assert ll == orc.smc_run(orc.MODEL_LG, th, y, N, seed, 0.5,
                                 orc.SYSTEMATIC, nthreads=4)


def test_residual_counts_closed_form():
    """Residual steps of the oracle sampler keep floor(N W_i) copies of each
    particle: with two particles of equal weight (N = 2) every step keeps
    one copy of each, so the sampler never moves a particle (x_t follows
    its own path) -- checked through the ESS history, which stays at N for
    the equal-weight steps a flat likelihood gives."""
    y = np.zeros(6)
    # sy huge: the likelihood is flat, the weights stay equal, ESS = N
    for scheme in (orc.RESIDUAL, orc.RESIDUAL_STRATIFIED,
                   orc.RESIDUAL_SYSTEMATIC):
        ll, ess, res = orc.smc_run(orc.MODEL_LG, (0.9, 1.0, 1e6), y, 2, 3,
                                   1.01, scheme, history=True)
        assert res.sum() == 6
        assert np.allclose(ess, 2.0, rtol=1e-9)
