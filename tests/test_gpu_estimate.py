"""SMCEstimator on the device (algorithm/smc.hpp:48-136).

The weighted-sum primitive is checked against the exact sum rounded once
(each product split exactly into four double products, math.fsum), which
is what its double-double accumulation promises; the reference's dgemv is
an external BLAS routine with no fixed summation order (none is built
here), so its results agree with these to its own rounding only.  The
sampler's selection estimates of x and x^2 are checked against the
oracle's sampler (same particles and weights; sums in another order) and,
on the linear-Gaussian model, against the Kalman filter's moments."""
import math
import os

import numpy as np
import pytest

import kalman
import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROW, COL = 101, 102


def _m():
    import mckl_amd as m
    return m


def _split(a):
    """Dekker split: a = hi + lo with 26-bit halves (exact)."""
    c = a * 134217729.0
    hi = c - (c - a)
    return hi, a - hi


def exact_weighted_sums(u, w):
    """sum_i w_i u_ij for every column j, the exact value rounded once."""
    wh, wl = _split(w)
    out = []
    for j in range(u.shape[1]):
        uh, ul = _split(u[:, j])
        terms = np.concatenate([wh * uh, wh * ul, wl * uh, wl * ul])
        out.append(math.fsum(terms.tolist()))
    return np.array(out)


@pytest.mark.parametrize("layout", [ROW, COL])
@pytest.mark.parametrize("udt", [np.float64, np.float32])
def test_weighted_sums_exact(dev, layout, udt):
    import torch
    m = _m()
    rng = np.random.default_rng(4)
    for n, dim in ((1, 1), (7, 3), (2048, 8), (2049, 9), (100003, 2),
                   (5000, 17), (0, 2), ((1 << 20) + 5, 3)):
        u = rng.normal(size=(n, dim)) * rng.choice([1e-3, 1.0, 1e4], (n, dim))
        u = u.astype(udt)
        w = rng.random(n)
        w /= max(w.sum(), 1.0)
        flat = u if layout == ROW else np.ascontiguousarray(u.T)
        got = m.estimate(torch.from_numpy(flat.reshape(-1).copy()).cuda(),
                         torch.from_numpy(w).cuda(), dim, layout)
        got = got.cpu().numpy()
        want = exact_weighted_sums(u.astype(np.float64), w)
        ulp = np.abs(got - want) / np.spacing(np.maximum(np.abs(want),
                                                         1e-300))
        assert np.all(ulp <= 1.0), (n, dim, ulp.max())
        # and the plain dot product (a dgemv's result) to its rounding
        ref = w @ u.astype(np.float64) if n else np.zeros(dim)
        bound = 4 * max(n, 1) * np.finfo(float).eps * (
            np.abs(w) @ np.abs(u.astype(np.float64)) if n else np.zeros(dim))
        assert np.all(np.abs(got - ref) <= bound + 1e-300)


def test_weighted_sums_errors(dev):
    import torch
    m = _m()
    u = torch.zeros(8, dtype=torch.float64, device="cuda")
    w = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(m.MCKLError):
        m.estimate(u, w, 2, 7)
    with pytest.raises(m.MCKLError):
        m.estimate(u, w, 3)            # u holds fewer than n * dim values


def _obs():
    return np.load(os.path.join(GOLDEN, "observations.npz"))


@pytest.mark.parametrize("scheme", ["SYSTEMATIC", "STRATIFIED",
                                    "MULTINOMIAL"])
def test_sampler_moments_vs_oracle(dev, scheme):
    m = _m()
    y = _obs()["lg_y"][:30]
    th = (0.9, 1.0, 0.5)
    sch = getattr(m, scheme)
    for N in (1, 1000, 5000):
        s = m.SMCSampler(m.MODEL_LG, th, y, N, 7, 0.5, sch)
        s.estimate_moments(3)
        s.iterate()
        est = s.estimates()
        ll, ess, res = s.results()
        oll, oess, ores, oest = orc.smc_run(orc.MODEL_LG, th, y, N, 7, 0.5,
                                            getattr(orc, scheme),
                                            nthreads=4, moments=3)
        assert np.array_equal(res, ores)
        assert abs(ll - oll) < 1e-9 * max(1.0, abs(oll))
        # same particles and weights; the oracle sums W_i x_i^j in order
        scale = np.abs(oest) + 1.0
        assert np.all(np.abs(est - oest) < 1e-9 * scale), (N, scheme)
        if N == 1:
            # one particle: W = 1, the estimates are its powers
            assert np.allclose(est[:, 1], est[:, 0] ** 2, rtol=1e-6)


def test_sampler_moments_vs_kalman(dev):
    """N = 2^20 particles: filtering means and second moments of every step
    within Monte Carlo error of the Kalman filter; turning the estimator off
    leaves the log-likelihood unchanged."""
    m = _m()
    y = _obs()["lg_y"]
    th = (0.9, 1.0, 0.5)
    km, kP = kalman.filter_moments(y, *th)
    s = m.SMCSampler(m.MODEL_LG, th, y, 1 << 20, 3, 0.5, m.SYSTEMATIC)
    s.estimate_moments(2)
    s.iterate()
    est = s.estimates()
    ll, ess, res = s.results()
    se = np.sqrt(kP / ess)
    z = (est[:, 0] - km) / se
    assert np.all(np.abs(z) < 6), np.abs(z).max()
    assert abs(z.mean()) < 6 / math.sqrt(len(z)) * 3
    assert np.all(np.abs(est[:, 1] - (kP + km * km)) <
                  6 * se * (2 * np.abs(km) + 2 * np.sqrt(kP)) + 1e-3)
    s.estimate_moments(0)
    s.iterate()
    assert s.log_likelihood() == ll
    with pytest.raises(m.MCKLError):
        s.estimates()
    with pytest.raises(m.MCKLError):
        s.estimate_moments(5)
