"""Synthetic workloads of the benchmark configurations, generated on the GPU.

Each generator uses only this package's fills (Philox4x32 + bulk Normal or
U01), so the inputs are reproducible bit for bit from their seeds; the
observation recursions run on the host in float64.

  philox_fill     Philox4x32, key 0x0123456789ABCDEF split little-endian into
                  key_type {0x89ABCDEF, 0x01234567}, counter 0.
  normal_streams  seeds 1..64, key {s, 0, ...}, counter 0, 2^22 samples each.
  lg_observations linear-Gaussian, x' = 0.9 x + e, y = x + 0.5 n, x0 ~ N(0,1),
                  z = rand(Philox4x32(seed), Normal<double>(0,1), 2T+1):
                  x0 = z0, x_t = a x_{t-1} + sx z_{2t-1}, y_t = x_t + sy z_{2t}.
  sv_observations stochastic volatility with the same draw layout:
                  x0 = s/sqrt(1-phi^2) z0, x_t = phi x_{t-1} + s z_{2t-1},
                  y_t = beta exp(x_t/2) z_{2t}.
  sv_batch        filter f: Philox4x32 key {f, 0xB47C4}; four U01OO<double>
                  draws u0..u3 give phi = 0.90+0.09u0, s = 0.1+0.2u1,
                  beta = 0.5+0.5u2, N_f = floor(exp(log 256 + u3 log 256))
                  (log-uniform on [256, 65536]); then 2T+1 Normal<double>
                  draws as in sv_observations.  SMC seed of filter f is f+1.
"""
import math

import numpy as np

PHILOX_KEY = (0x89ABCDEF, 0x01234567)


def _torch():
    import torch
    return torch


def normal_f64(seed, n):
    import mckl_amd as m
    torch = _torch()
    rng = m.Philox4x32(key=[seed, 0])
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    m.rand(rng, m.NormalDistribution(0.0, 1.0, "float64"), n, out)
    return out.cpu().numpy()


def lg_from_z(z, T, a=0.9, sx=1.0, sy=0.5):
    x = z[0]
    y = np.empty(T)
    for t in range(1, T + 1):
        x = a * x + sx * z[2 * t - 1]
        y[t - 1] = x + sy * z[2 * t]
    return y


def sv_from_z(z, T, phi, s, beta):
    x = s / math.sqrt(1.0 - phi * phi) * z[0]
    y = np.empty(T)
    for t in range(1, T + 1):
        x = phi * x + s * z[2 * t - 1]
        y[t - 1] = beta * math.exp(x / 2.0) * z[2 * t]
    return y


def lg_observations(T=100, seed=7, theta=(0.9, 1.0, 0.5)):
    return lg_from_z(normal_f64(seed, 2 * T + 1), T, *theta)


def sv_observations(T=500, seed=11, theta=(0.95, 0.2, 0.7)):
    return sv_from_z(normal_f64(seed, 2 * T + 1), T, *theta)


def sv_params_from_u(u):
    phi = 0.90 + 0.09 * u[0]
    s = 0.1 + 0.2 * u[1]
    beta = 0.5 + 0.5 * u[2]
    nf = int(math.floor(math.exp(math.log(256.0) + u[3] * math.log(256.0))))
    return phi, s, beta, min(max(nf, 256), 65536)


def sv_batch(nfilters=4096, T=50):
    """theta (nfilters, 3), N (nfilters,) u32, y (nfilters, T), seeds."""
    import mckl_amd as m
    torch = _torch()
    theta = np.empty((nfilters, 3))
    N = np.empty(nfilters, np.uint32)
    y = np.empty((nfilters, T))
    u = torch.empty(4, dtype=torch.float64, device="cuda")
    z = torch.empty(2 * T + 1, dtype=torch.float64, device="cuda")
    for f in range(nfilters):
        rng = m.Philox4x32(key=[f, 0xB47C4])
        m.rand(rng, m.U01OODistribution("float64"), 4, u)
        m.rand(rng, m.NormalDistribution(0.0, 1.0, "float64"), 2 * T + 1, z)
        uh = u.cpu().numpy()
        phi, s, beta, nf = sv_params_from_u(uh)
        theta[f] = (phi, s, beta)
        N[f] = nf
        y[f] = sv_from_z(z.cpu().numpy(), T, phi, s, beta)
    seeds = np.arange(1, nfilters + 1, dtype=np.uint64)
    return theta, N, y, seeds


def sv_grad_points(nb=256, theta=(0.95, 0.2, 0.7), rel=0.02):
    """Gradient batch: nb parameter points jittered around theta,
    theta_p,j = theta_j (1 + rel (u_{3p+j} - 1/2)) with u the U01OO<double>
    stream of Philox4x32 key {0x6D1A7, 0}, counter 0."""
    import mckl_amd as m
    torch = _torch()
    rng = m.Philox4x32(key=[0x6D1A7, 0])
    u = torch.empty(3 * nb, dtype=torch.float64, device="cuda")
    m.rand(rng, m.U01OODistribution("float64"), 3 * nb, u)
    u = u.cpu().numpy().reshape(nb, 3)
    return np.asarray(theta, np.float64) * (1.0 + rel * (u - 0.5))
