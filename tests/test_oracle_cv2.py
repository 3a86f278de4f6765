"""The oracle's two-dimensional model (model 4, constant velocity: position
p_t = p + v + sp e0, velocity v_t = v + sv e1, y_t = p_t + sy w, both
components starting N(0, 1); the model the fused-sampler tests run with a
vector state) against an exact Kalman filter written here:

* the Kalman log-likelihood equals the dense Gaussian density of y_{1:T}
  (its covariance built from the model's linear recursion);
* over many seeds the oracle's log-likelihood estimates agree with it
  within Monte Carlo error, and a model with the velocity noise doubled is
  told apart.
"""
import math

import numpy as np

import oracle as orc

TH = (0.3, 0.2, 0.8)  # (sp, sv, sy)
A = np.array([[1.0, 1.0], [0.0, 1.0]])


def simulate(T, seed):
    rng = np.random.default_rng(seed)
    sp, sv, sy = TH
    x = rng.standard_normal(2)
    y = np.zeros(T)
    for t in range(T):
        x = A @ x + np.array([sp, sv]) * rng.standard_normal(2)
        y[t] = x[0] + sy * rng.standard_normal()
    return y


def kalman(y, th):
    sp, sv, sy = th
    Q = np.diag([sp * sp, sv * sv])
    m = np.zeros(2)
    P = np.eye(2)
    ll = 0.0
    for yt in y:
        m = A @ m
        P = A @ P @ A.T + Q
        S = P[0, 0] + sy * sy
        r = yt - m[0]
        ll += -0.5 * (math.log(2 * math.pi * S) + r * r / S)
        K = P[:, 0] / S
        m = m + K * r
        P = P - np.outer(K, P[0, :])
    return ll


def dense(y, th):
    """log N(y; 0, C) with C from x_t = A^t x_0 + sum_s A^(t-s) q_s."""
    sp, sv, sy = th
    T = len(y)
    Q = np.diag([sp * sp, sv * sv])
    # Cov(x_s, x_t) for s <= t: A^(t-s) Cov(x_s)
    covx = []
    P = np.eye(2)
    for _ in range(T):
        P = A @ P @ A.T + Q
        covx.append(P.copy())
    C = np.zeros((T, T))
    for s in range(T):
        for t in range(s, T):
            c = np.linalg.matrix_power(A, t - s) @ covx[s]
            C[s, t] = C[t, s] = c[0, 0]
    C += sy * sy * np.eye(T)
    sign, logdet = np.linalg.slogdet(C)
    return -0.5 * (T * math.log(2 * math.pi) + logdet + y @ np.linalg.solve(C, y))


def test_kalman_matches_dense_gaussian():
    y = simulate(12, 3)
    assert abs(kalman(y, TH) - dense(y, TH)) < 1e-10 * abs(dense(y, TH))
    th2 = (0.5, 0.05, 1.3)
    assert abs(kalman(y, th2) - dense(y, th2)) < 1e-10 * abs(dense(y, th2))


def test_cv2_oracle_vs_kalman():
    T, N, reps = 25, 1 << 13, 24
    y = simulate(T, 5)
    exact = kalman(y, TH)
    a = np.array([orc.smc_run(4, TH, y, N, 100 + r, 0.5, orc.SYSTEMATIC,
                              nthreads=8) for r in range(reps)])
    se = a.std(ddof=1) / math.sqrt(reps)
    print("oracle %.5f +- %.5f, Kalman %.5f" % (a.mean(), se, exact))
    assert abs(a.mean() - exact) < 4 * max(se, 1e-4)
    # the velocity noise doubled is told apart
    bad = (TH[0], 2 * TH[1], TH[2])
    b = np.array([orc.smc_run(4, bad, y, N, 100 + r, 0.5, orc.SYSTEMATIC,
                              nthreads=8) for r in range(6)])
    assert abs(b.mean() - exact) > 10 * max(se, 1e-4)
    assert abs(kalman(y, bad) - b.mean()) < 6 * max(b.std(ddof=1), 1e-4)
