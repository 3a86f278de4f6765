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


def filter_moments(y, a, sx, sy, m0=0.0, v0=1.0):
    """Filtering means E[x_t | y_1..y_t] and variances, t = 1..T, by the
    predict/update recursion of ``loglik``."""
    m, P = m0, v0
    ms, Ps = [], []
    for yt in y:
        m = a * m
        P = a * a * P + sx * sx
        S = P + sy * sy
        K = P / S
        m = m + K * (yt - m)
        P = (1.0 - K) * P
        ms.append(m)
        Ps.append(P)
    return np.array(ms), np.array(Ps)


def filter_moments_dense(y, a, sx, sy, m0=0.0, v0=1.0):
    """The same moments by conditioning the joint Gaussian of (x_t, y_1..y_t)
    directly: E = mu_x + c' C^-1 (y - mu_y), Var = var_x - c' C^-1 c."""
    T = len(y)
    var = np.empty(T + 1)
    var[0] = v0
    for s in range(1, T + 1):
        var[s] = a * a * var[s - 1] + sx * sx

    def cov(s, t):  # Cov(x_s, x_t)
        s, t = min(s, t), max(s, t)
        return a ** (t - s) * var[s]

    ms, Ps = [], []
    for t in range(1, T + 1):
        C = np.array([[cov(i, j) for j in range(1, t + 1)]
                      for i in range(1, t + 1)]) + np.eye(t) * sy * sy
        c = np.array([cov(t, j) for j in range(1, t + 1)])
        mu_y = np.array([a ** j * m0 for j in range(1, t + 1)])
        sol = np.linalg.solve(C, np.asarray(y[:t]) - mu_y)
        ms.append(a ** t * m0 + c @ sol)
        Ps.append(var[t] - c @ np.linalg.solve(C, c))
    return np.array(ms), np.array(Ps)
