"""The oracle's growth model (model 3: x_t = (0.5 x + 25 x / (1 + x^2)) +
u_t + s e, y_t = x^2 / 20 + sy w, the model the fused-sampler tests run as a
caller's own) against an independent FP64 bootstrap filter written here with
numpy's generator: over several seeds the two log-likelihood estimates agree
within Monte Carlo error.  A dropped drift, a wrong sign or a transposed
parameter in the oracle's restatement moves its estimate by many standard
errors."""
import math

import numpy as np

import oracle as orc

GROWTH = (math.sqrt(10.0), 1.0, math.sqrt(5.0))  # (s, sy, sd0)


def data(T, seed):
    rng = np.random.default_rng(seed)
    s, sy, sd0 = GROWTH
    x = sd0 * rng.standard_normal()
    y = np.zeros(T)
    u = 8.0 * np.cos(1.2 * np.arange(1, T + 1))
    for t in range(T):
        x = 0.5 * x + 25.0 * x / (1.0 + x * x) + u[t] + s * rng.standard_normal()
        y[t] = x * x / 20.0 + sy * rng.standard_normal()
    return y, u


def numpy_pf(y, u, N, rng):
    """Bootstrap filter, multinomial resampling at every step."""
    s, sy, sd0 = GROWTH
    c = -math.log(sy) - 0.5 * math.log(2.0 * math.pi)
    x = sd0 * rng.standard_normal(N)
    ll = 0.0
    for t in range(len(y)):
        x = 0.5 * x + 25.0 * x / (1.0 + x * x) + u[t] + \
            s * rng.standard_normal(N)
        lw = -0.5 * ((y[t] - x * x / 20.0) / sy) ** 2 + c
        m = lw.max()
        w = np.exp(lw - m)
        ll += m + math.log(w.mean())
        x = x[rng.choice(N, N, p=w / w.sum())]
    return ll


def test_growth_oracle_vs_numpy_filter():
    T, N, reps = 20, 1 << 14, 12
    y, u = data(T, 5)
    yy = np.concatenate([y, u])
    a = np.array([orc.smc_run(3, GROWTH, yy, N, 100 + r, 1.0, 0, nthreads=8)
                  for r in range(reps)])
    rng = np.random.default_rng(77)
    b = np.array([numpy_pf(y, u, N, rng) for _ in range(reps)])
    se = math.sqrt(a.var(ddof=1) / reps + b.var(ddof=1) / reps)
    print("oracle %.4f +- %.4f, numpy %.4f +- %.4f" % (
        a.mean(), a.std(ddof=1), b.mean(), b.std(ddof=1)))
    assert abs(a.mean() - b.mean()) < 5 * max(se, 1e-3)
    # and a model that differs in one term is told apart (half the drift;
    # negating it would be a symmetry, x -> -x)
    bad = np.concatenate([y, 0.5 * u])
    c = np.array([orc.smc_run(3, GROWTH, bad, N, 100 + r, 1.0, 0, nthreads=8)
                  for r in range(4)])
    assert abs(c.mean() - b.mean()) > 20 * max(se, 1e-3)
