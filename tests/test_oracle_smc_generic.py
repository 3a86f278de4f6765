"""Pin the oracle's SMCSampler<T> restatement (GenericSMC: do_iterate's
order, Particle::select, ResampleEval with Particle::rng()) against the
reference's own SMCSampler (oracle/_ref/libmckl_ref_smc.so, built from the
reference headers) on a linear-Gaussian model whose noise comes from a
table: ESS and size histories and the final states agree bit for bit for
every scheme and the threshold's special values.  CPU only."""
import os

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

needs_ref = pytest.mark.skipif(orc.REF_SMC is None,
                               reason="reference SMCSampler driver not built")


def _case(N, T, seed):
    rng = np.random.default_rng(seed)
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"][:T]
    z = rng.normal(size=T * N)
    return y, z


def _oracle(N, T, scheme, thr, key, a, sx, sy, mut, y, z):
    smc = orc.GenericSMC(N, 1, key)
    orc.lg_table_model(smc, y, z, a, sx, sy, mut)
    smc.evals[1].append(orc.resample_eval(scheme))
    smc.threshold = thr
    smc.iterate(T)
    return (np.array(smc.ess_history), np.array(smc.size_history),
            smc.particle.x[:, 0])


@needs_ref
@pytest.mark.parametrize("scheme", range(6))
@pytest.mark.parametrize("N,thr,mut", [(1000, 0.5, 1.0), (7, np.inf, 0.999),
                                       (1, 0.5, 1.0), (513, -np.inf, 1.0),
                                       (2000, 0.9, 0.999)])
def test_generic_oracle_matches_reference(scheme, N, thr, mut):
    T = 25
    y, z = _case(N, T, 100 + N + scheme)
    key = (12345 + scheme, 0xABCDEF)
    args = (N, T, scheme, thr, key, 0.9, 1.0, 0.5, mut, y, z)
    ess, size, x = _oracle(*args)
    ress, rsize, rx = orc.ref_smc_lg(*args)
    assert np.array_equal(size, rsize)
    assert np.array_equal(ess, ress)
    assert np.array_equal(x, rx)
    if thr == np.inf and N > 1:
        # resampling every step: the ESS after a resample-and-move step is
        # below N whenever the weights are unequal
        assert np.all(ess[1:] < N)


def test_generic_oracle_never_resamples():
    """threshold = -inf (resample_threshold_never): no step resamples, so
    the weights accumulate every increment and the state is the plain
    recursion of the table noise."""
    N, T = 50, 10
    y, z = _case(N, T, 3)
    ess, size, x = _oracle(N, T, orc.SYSTEMATIC, -np.inf, (1, 0), 0.9, 1.0,
                           0.5, 1.0, y, z)
    xs = z[:N].copy()
    for t in range(1, T):
        xs = 0.9 * xs + 1.0 * z[t * N:(t + 1) * N]
    assert np.array_equal(x, xs)
    # the ESS of the accumulated weights from scratch
    lw = np.zeros(N)
    xs = z[:N].copy()
    for t in range(T):
        if t:
            xs = 0.9 * xs + 1.0 * z[t * N:(t + 1) * N]
        d = (y[t] - xs) / 0.5
        lw += -0.5 * d * d
        w = np.exp(lw - lw.max())
        w /= w.sum()
        assert abs(ess[t] - 1.0 / (w * w).sum()) < 1e-9 * N


@needs_ref
@pytest.mark.parametrize("scheme,thr", [(orc.SYSTEMATIC, 0.5),
                                        (orc.STRATIFIED, 0.8),
                                        (orc.RESIDUAL, np.inf)])
def test_generic_pmmh_oracle_matches_reference(scheme, thr):
    """The oracle's PMCMCMutation restatement over GenericSMC against the
    reference's own PMCMCMutation<Param, PMCMCStateMatrix<...>> on the
    table-noise model: acceptance flags, parameters and log normalising
    constants bit for bit over 25 iterations (including the particles a
    rejected step leaves behind)."""
    N, T, iters = 300, 12, 25
    y, z = _case(N, T, 40 + scheme)
    args = (N, T, iters, scheme, thr, (777, scheme), (0.08, 0.08, 0.08),
            (0.7, 0.9, 0.6), y, z)
    acc, par, lnc = orc.generic_pmmh_lg(*args)
    racc, rpar, rlnc = orc.ref_pmcmc_lg(*args)
    assert np.array_equal(acc, racc)
    assert np.array_equal(par, rpar)
    assert np.array_equal(lnc, rlnc)
    assert 0 < acc.sum() < iters - 1
