"""Kalman-filter oracle for the linear-Gaussian SMC configuration (test
infrastructure only).

Model: x_0 ~ N(m0, v0), x_t = a x_{t-1} + sx e_t, y_t = x_t + sy n_t,
observations y_1..y_T.  ``loglik`` is the textbook predict/update recursion
in float64; ``loglik_dense`` evaluates the same likelihood as one
multivariate normal density with the covariance of (y_1..y_T) written in
closed form, which pins the recursion independently; ``grad`` returns the
exact gradient of ``loglik`` with respect to (a, sx, sy) by forward-mode
sensitivity recursions.
"""
import math

import numpy as np

LOG2PI = math.log(2.0 * math.pi)


def loglik(y, a, sx, sy, m0=0.0, v0=1.0):
    m, P = m0, v0
    ll = 0.0
    for yt in y:
        m = a * m
        P = a * a * P + sx * sx
        S = P + sy * sy
        r = yt - m
        ll += -0.5 * (LOG2PI + math.log(S) + r * r / S)
        K = P / S
        m = m + K * r
        P = (1.0 - K) * P
    return ll


def loglik_dense(y, a, sx, sy, m0=0.0, v0=1.0):
    T = len(y)
    # Cov(x_s, x_t) for s <= t: a^{t-s} Var(x_s), Var(x_s) = a^{2s} v0 +
    # sx^2 sum_{k<s} a^{2k}
    var = np.empty(T + 1)
    var[0] = v0
    for s in range(1, T + 1):
        var[s] = a * a * var[s - 1] + sx * sx
    C = np.empty((T, T))
    for s in range(1, T + 1):
        for t in range(s, T + 1):
            C[s - 1, t - 1] = C[t - 1, s - 1] = a ** (t - s) * var[s]
    C += np.eye(T) * sy * sy
    mu = np.array([a ** t * m0 for t in range(1, T + 1)])
    r = np.asarray(y) - mu
    sign, logdet = np.linalg.slogdet(C)
    return -0.5 * (T * LOG2PI + logdet + r @ np.linalg.solve(C, r))


def grad(y, a, sx, sy, m0=0.0, v0=1.0):
    """d loglik / d(a, sx, sy) by differentiating the recursion."""
    m, P = m0, v0
    dm = np.zeros(3)
    dP = np.zeros(3)
    g = np.zeros(3)
    for yt in y:
        # predict
        dm = np.array([m + a * dm[0], a * dm[1], a * dm[2]])
        m = a * m
        dP = np.array([2 * a * P + a * a * dP[0], a * a * dP[1] + 2 * sx,
                       a * a * dP[2]])
        P = a * a * P + sx * sx
        S = P + sy * sy
        dS = dP + np.array([0.0, 0.0, 2 * sy])
        r = yt - m
        dr = -dm
        g += -0.5 * (dS / S + 2 * r * dr / S - r * r * dS / (S * S))
        K = P / S
        dK = dP / S - P * dS / (S * S)
        dm = dm + dK * r + K * dr
        m = m + K * r
        dP = -dK * P + (1.0 - K) * dP
        P = (1.0 - K) * P
    return g
