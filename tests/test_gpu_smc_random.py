"""The fused samplers against the oracle over configurations drawn at
random (a fixed generator, so the cases are the same every run): particle
counts from 1 to about 300000 (tile boundaries, odd sizes), 1 to 25 steps,
thresholds from never to always resampling, every scheme, both built-in
models; and batches of random filters against their plain runs.  The
parity bar is the one of the named cases: the same resampling decisions,
ESS to 1e-11, log Z to 1e-9 relative; batched runs bit-identical."""
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
THETA = {1: (0.9, 1.0, 0.5), 2: (0.95, 0.2, 0.7)}


def _cases(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        model = int(rng.integers(1, 3))
        kind = rng.integers(0, 4)
        if kind == 0:  # near a tile boundary
            N = int(2048 * rng.integers(1, 40) + rng.integers(-2, 3))
        elif kind == 1:  # tiny
            N = int(rng.integers(1, 40))
        else:  # log-uniform
            N = int(np.exp(rng.uniform(np.log(50), np.log(300000))))
        N = max(N, 1)
        T = int(rng.integers(1, 26))
        thr = float(rng.choice([0.0, 0.3, 0.5, 0.9, 1.0, 1.01]))
        scheme = int(rng.integers(0, 6))
        out.append((model, N, T, thr, scheme, int(rng.integers(1, 1 << 30))))
    return out


@pytest.mark.parametrize("case", _cases(36, 2024))
def test_random_plain_vs_oracle(dev, case):
    import mckl_amd as m
    model, N, T, thr, scheme, seed = case
    ob = np.load(os.path.join(GOLDEN, "observations.npz"))
    y = (ob["lg_y"] if model == 1 else ob["sv_y"])[:T]
    th = THETA[model]
    s = m.SMCSampler(model, th, y, N, seed, thr, scheme)
    s.iterate()
    ll, ess, res = s.results()
    oll, oess, ores = orc.smc_run(model, th, y, N, seed, thr, scheme,
                                  nthreads=16, history=True)
    print(case, ll, oll, int(res.sum()))
    assert np.array_equal(res, ores)
    assert np.allclose(ess, oess, rtol=1e-11, atol=0)
    assert abs(ll - oll) < 1e-9 * max(1.0, abs(oll))


@pytest.mark.parametrize("seed", [7, 8, 9])
def test_random_batches_equal_plain(dev, seed):
    import mckl_amd as m
    rng = np.random.default_rng(seed)
    ob = np.load(os.path.join(GOLDEN, "observations.npz"))
    nf = int(rng.integers(3, 12))
    N = np.exp(rng.uniform(0, np.log(131072), nf)).astype(np.uint32)
    N = np.maximum(N, 1)
    T = int(rng.integers(1, 20))
    theta = ob["batch64_theta"][:nf]
    y = ob["batch64_y"][:nf, :T]
    seeds = rng.integers(1, 1 << 40, nf).astype(np.uint64)
    thr = float(rng.choice([0.3, 0.5, 1.01]))
    scheme = int(rng.integers(0, 6))
    b = m.SMCBatch(2, N, theta, y, seeds, thr, scheme)
    b.run()
    lb = b.log_likelihood()
    for f in range(nf):
        lp = m.smc_loglik(2, theta[f], y[f], int(N[f]), int(seeds[f]), thr,
                          scheme)
        assert lb[f] == lp, (seed, f, int(N[f]), scheme, lb[f], lp)
