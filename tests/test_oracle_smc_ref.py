"""The fused samplers' resampling step (the oracle's restatement, which the
device matches bit for bit) against the reference's own transforms,
compiled from its headers (oracle/_ref):

* the step's uniform sequence equals u01_trans_sorted / stratified /
  systematic (algorithm/resample.hpp:141-212) applied to the same uniform
  draws, bit for bit;
* the replication numbers equal resample_trans_u01_rep
  (algorithm/resample.hpp:374-407) on the step's normalised weights and
  that sequence, except where a uniform lies within rounding of a
  cumulative weight: the reference sums the normalised weights
  sequentially in FP64, the samplers round the exact prefix once.  Each
  difference must have a uniform within N eps relative of the exact
  cumulative weight at that particle; the count within 1e-15 is printed.

Models: linear-Gaussian and stochastic volatility (the built-in ones) and
the growth model (a caller's own in the fused-sampler tests), at several
steps and sizes, all six schemes (the residual ones through
resample_trans_residual, the integral parts exact).
"""
import math
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.skipif(orc.REF is None,
                                reason="oracle/_ref not built (no reference)")

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GROWTH = (math.sqrt(10.0), 1.0, math.sqrt(5.0))


def growth_data(T, seed):
    rng = np.random.default_rng(seed)
    s, sy, sd0 = GROWTH
    x = sd0 * rng.standard_normal()
    y = np.zeros(T)
    u = 8.0 * np.cos(1.2 * np.arange(1, T + 1))
    for t in range(T):
        x = 0.5 * x + 25.0 * x / (1.0 + x * x) + u[t] + s * rng.standard_normal()
        y[t] = x * x / 20.0 + sy * rng.standard_normal()
    return np.concatenate([y, u])


def model_data(model):
    ob = np.load(os.path.join(GOLDEN, "observations.npz"))
    if model == 1:
        return (0.9, 1.0, 0.5), ob["lg_y"][:40]
    if model == 2:
        return (0.95, 0.2, 0.7), ob["sv_y"][:40]
    return GROWTH, growth_data(40, 21)


def unexplained(W, u, rep_ref, rep):
    """(differences, of them within 1e-15 relative); asserts each is a
    uniform within N eps relative of an exact cumulative weight."""
    diff = np.nonzero(rep_ref.astype(np.int64) != rep.astype(np.int64))[0]
    if len(diff) == 0:
        return 0, 0
    exact = np.cumsum(np.asarray(W, np.longdouble))
    N = len(W)
    tol = N * np.finfo(np.float64).eps
    within = 0
    for i in diff:
        d = 1.0
        for k in (i - 1, i):
            if 0 <= k < N - 1:
                S = float(exact[k])
                d = min(d, float(np.min(np.abs(u - S))) / S)
        assert d <= tol, (i, d)
        within += d <= 1e-15
    return len(diff), within


@pytest.mark.parametrize("model", [1, 2, 3])
@pytest.mark.parametrize("scheme", [orc.MULTINOMIAL, orc.STRATIFIED,
                                    orc.SYSTEMATIC])
def test_resampling_step_vs_reference(model, scheme):
    th, y = model_data(model)
    total = within = steps = 0
    for N, seed, t_first in ((5000, 6, 1), (70001, 7, 4), (2049, 8, 9)):
        d = orc.smc_resample_dump(model, th, y, N, seed, 0.5, scheme, t_first)
        assert d is not None
        t, u01, W, u, rep = d
        ru = orc.ref_u01_trans(scheme, u01, N)
        assert np.array_equal(ru, u), (N, t)
        rr = orc.ref_trans_u01_rep(W, ru)
        assert int(rep.sum()) == N and int(rr.sum()) == N
        nd, n15 = unexplained(W, u, rr, rep)
        total += nd
        within += n15
        steps += 1
    print("model %d scheme %d: %d steps, rep differences %d (within 1e-15: "
          "%d)" % (model, scheme, steps, total, within))


@pytest.mark.parametrize("model", [1, 2, 3])
@pytest.mark.parametrize("scheme", [orc.RESIDUAL, orc.RESIDUAL_STRATIFIED,
                                    orc.RESIDUAL_SYSTEMATIC])
def test_residual_step_vs_reference(model, scheme):
    """Residual schemes (algorithm/resample.hpp:333-364, 557-574): the
    integral parts of N W_i equal resample_trans_residual's, the R uniforms
    equal the base scheme's transform of the same draws, and the
    replication numbers equal resample_trans_u01_rep on the reference's
    normalised residuals plus the integral parts, except at rounding-edge
    uniforms of the residuals' cumulative sums."""
    th, y = model_data(model)
    base = scheme - orc.RESIDUAL
    total = within = 0
    for N, seed, t_first in ((5000, 6, 1), (70001, 7, 4), (2049, 8, 9)):
        d = orc.smc_resample_dump(model, th, y, N, seed, 0.5, scheme, t_first)
        assert d is not None
        t, u01, W, u, rep = d
        resid, integ, R = orc.ref_trans_residual(W)
        assert R == len(u) and R > 0
        ru = orc.ref_u01_trans(base, u01, R)
        assert np.array_equal(ru, u), (N, t)
        rr = orc.ref_trans_u01_rep(resid, ru) + integ
        assert int(rep.sum()) == N and int(rr.sum()) == N
        nd, n15 = unexplained(resid, u, rr, rep)
        total += nd
        within += n15
    print("model %d scheme %d: rep differences %d (within 1e-15: %d)"
          % (model, scheme, total, within))


def test_resampling_step_large_vs_reference():
    """2^20 + 3 particles: the sizes where the two cumulative sums drift
    furthest apart."""
    th, y = model_data(1)
    N = (1 << 20) + 3
    for scheme in (orc.SYSTEMATIC, orc.STRATIFIED):
        t, u01, W, u, rep = orc.smc_resample_dump(1, th, y[:8], N, 31, 0.99,
                                                  scheme, 2)
        ru = orc.ref_u01_trans(scheme, u01, N)
        assert np.array_equal(ru, u)
        rr = orc.ref_trans_u01_rep(W, ru)
        nd, n15 = unexplained(W, u, rr, rep)
        print("scheme %d N=%d step %d: rep differences %d (within 1e-15: %d)"
              % (scheme, N, t, nd, n15))


def test_dump_matches_the_run():
    """The dump is the run's own step: the same decisions as smc_run up to
    it, and a planted change in the weights is seen by the reference's
    transform (the check is not vacuous)."""
    th, y = model_data(2)
    N = 4096
    _, _, res = orc.smc_run(2, th, y, N, 5, 0.5, orc.SYSTEMATIC,
                            history=True)
    t, u01, W, u, rep = orc.smc_resample_dump(2, th, y, N, 5, 0.5,
                                              orc.SYSTEMATIC)
    assert res[t - 1] == 1 and not res[:t - 1].any()
    # move half of one particle's weight to its neighbour: the reference's
    # counts must then differ from the step's own beyond any rounding
    W2 = W.copy()
    i = int(np.argmax(W2[:-1]))
    W2[i + 1] += W2[i] / 2
    W2[i] /= 2
    rr = orc.ref_trans_u01_rep(W2, u)
    with pytest.raises(AssertionError):
        unexplained(W, u, rr, rep)
