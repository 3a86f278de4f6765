"""SMCSampler<T> with evaluation objects on device particles
(mckl_amd.GenericSMCSampler) against the oracle's restatement, which is
pinned bit for bit to the reference's own SMCSampler on the same model
(tests/test_oracle_smc_generic.py).

The model is the linear-Gaussian one with its noise from a table, written
twice as user evaluation objects: torch operations on the device state
here, numpy operations in the oracle, one rounding per operation in both.
The device weights agree with the oracle's to ~1e-13 (another summation
order), so the ESS histories agree to that and the resampling decisions
and ancestors are the same; the states then agree exactly."""
import math
import os

import numpy as np
import pytest

import kalman
import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _m():
    import mckl_amd as m
    return m


def _case(N, T, seed):
    rng = np.random.default_rng(seed)
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"][:T]
    z = rng.normal(size=T * N)
    return y, z


def _gpu_lg(s, y, z, a, sx, sy, mut):
    import torch
    N = s.size()
    zd = torch.from_numpy(z).cuda()

    def sel(t, p):
        x = p.state[:, 0]
        xn = zd[:N].clone() if t == 0 else x * a + zd[t * N:(t + 1) * N] * sx
        p.state[:, 0] = xn
        d = (float(y[t]) - xn) / sy
        p.weight().add_log((d * d) * -0.5)

    s.selection(sel)
    if mut != 1.0:
        def mu(t, p):
            p.state.mul_(mut)
        s.mutation(mu)


def _oracle(N, T, scheme, thr, key, a, sx, sy, mut, y, z):
    smc = orc.GenericSMC(N, 1, key)
    orc.lg_table_model(smc, y, z, a, sx, sy, mut)
    smc.evals[1].append(orc.resample_eval(scheme))
    smc.threshold = thr
    smc.iterate(T)
    return smc


@pytest.mark.parametrize("scheme", range(6))
@pytest.mark.parametrize("N,thr,mut", [(1000, 0.5, 1.0), (7, math.inf, 0.999),
                                       (1, 0.5, 1.0), (513, -math.inf, 1.0),
                                       (5000, 0.9, 0.999)])
def test_generic_matches_oracle(dev, scheme, N, thr, mut):
    m = _m()
    T = 25
    y, z = _case(N, T, 100 + N + scheme)
    key = (12345 + scheme, 0xABCDEF)
    s = m.GenericSMCSampler(N, 1, rng=m.Philox4x32(key=key))
    _gpu_lg(s, y, z, 0.9, 1.0, 0.5, mut)
    s.resample(scheme)
    s.resample_threshold(thr)
    s.iterate(T)
    o = _oracle(N, T, scheme, thr, key, 0.9, 1.0, 0.5, mut, y, z)
    assert s.num_iter() == T
    assert np.array_equal(s.size_history(), np.array(o.size_history))
    oess = np.array(o.ess_history)
    assert np.allclose(s.ess_history(), oess, rtol=1e-11, atol=0)
    # same decisions, same ancestors: the states are the same numbers
    x = s.particle().state[:, 0].cpu().numpy()
    assert np.array_equal(x, o.particle.x[:, 0])
    assert np.allclose(s.particle().weight().data.cpu().numpy(),
                       o.particle.W, rtol=1e-12, atol=0)


def _split(a):
    c = a * 134217729.0
    hi = c - (c - a)
    return hi, a - hi


def _exact(u, w):
    wh, wl = _split(w)
    out = []
    for j in range(u.shape[1]):
        uh, ul = _split(u[:, j])
        out.append(math.fsum(np.concatenate(
            [wh * uh, wh * ul, wl * uh, wl * ul]).tolist()))
    return np.array(out)


def test_generic_estimators_and_order(dev):
    """Estimators in all three slots (row-major, column-major, record-only)
    against the exactly rounded weighted sums of the oracle run's particles,
    the evaluation order of do_iterate, and summary()."""
    import torch
    m = _m()
    N, T = 3000, 12
    y, z = _case(N, T, 7)
    key = (99, 7)
    log = []
    s = m.GenericSMCSampler(N, 1, rng=m.Philox4x32(key=key))
    _gpu_lg(s, y, z, 0.9, 1.0, 0.5, 1.0)
    s.selection(lambda t, p: log.append(("sel", t)))
    s.resample(m.STRATIFIED)
    s.resample(lambda t, p: log.append(("res", t)))
    s.mutation(lambda t, p: log.append(("mut", t)))

    def moments(t, d, p, u):                 # u: (N, 2)
        x = p.state[:, 0].to(torch.float64)
        u[:, 0] = x
        u[:, 1] = x * x

    def moments_col(t, d, p, u):             # u: (2, N)
        x = p.state[:, 0].to(torch.float64)
        u[0] = x
        u[1] = x * x

    def ess_rec(t, d, p, r):
        r[0] = p.weight().ess()

    e0 = m.SMCEstimator(2, moments)
    e1 = m.SMCEstimator(2, moments_col, m.ColMajor)
    e2 = m.SMCEstimator(1, ess_rec, record_only=True)
    assert s.selection_estimator(e0) == 0
    assert s.resample_estimator(e1) == 0
    assert s.mutation_estimator(e2) == 0
    assert s.selection_estimator(0) is e0
    s.resample_threshold(0.4)
    s.iterate(T)
    o = orc.GenericSMC(N, 1, key)
    orc.lg_table_model(o, y, z, 0.9, 1.0, 0.5)
    o.evals[1].append(orc.resample_eval(orc.STRATIFIED))
    o.threshold = 0.4
    o.estimators[0].append(lambda t, p: _exact(
        np.stack([p.x[:, 0], p.x[:, 0] ** 2], 1), p.W))
    o.estimators[1].append(lambda t, p: _exact(
        np.stack([p.x[:, 0], p.x[:, 0] ** 2], 1), p.W))
    o.iterate(T)
    got0, got1 = e0.estimates(), e1.estimates()
    want0, want1 = np.array(o.estimates[0]), np.array(o.estimates[1])
    # same particles; the weights agree to ~1e-12 relative
    assert np.all(np.abs(got0 - want0) <= 1e-11 * (np.abs(want0) + 1.0))
    assert np.all(np.abs(got1 - want1) <= 1e-11 * (np.abs(want1) + 1.0))
    ess = s.ess_history()
    res = ess < 0.4 * N
    # record-only: the ESS after the mutation step, N after a resample
    assert np.allclose(e2.estimates()[:, 0], np.where(res, float(N), ess),
                       rtol=1e-12, atol=0)
    want_log = []
    for t in range(T):
        want_log.append(("sel", t))
        if res[t]:
            want_log.append(("res", t))
        want_log.append(("mut", t))
    assert log == want_log
    assert 0 < res.sum() < T
    sm = s.summary()
    assert sm.shape == (T, 5)
    assert np.array_equal(sm[:, :2], got0) and np.array_equal(sm[:, 2:4], got1)
    assert np.allclose(e0.average(), got0.mean(0))
    s.clear()
    assert s.num_iter() == 0 and e0.num_iter() == 0
    s.reset()
    s.iterate(1)
    assert s.num_iter() == 1 and len(log) == 3 * T - (T - res.sum())


def test_generic_kalman_f32(dev):
    """A float32 bootstrap filter written as evaluation objects (normals from
    the library's Normal fill): the sum of the log normaliser increments is
    within Monte Carlo error of the Kalman log-likelihood."""
    import torch
    m = _m()
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"][:50]
    a, sx, sy = 0.9, 1.0, 0.5
    N = 1 << 16
    s = m.GenericSMCSampler(N, 1, torch.float32, seed=5)
    eng = m.Philox4x32(11)
    z = torch.empty(N, dtype=torch.float32, device="cuda")
    dist = m.NormalDistribution(0.0, 1.0, "float32")
    c = -math.log(sy) - 0.5 * math.log(2 * math.pi)
    lnc = []

    def sel(t, p):
        if t == 0:                           # x_0 ~ N(0, 1)
            m.rand(eng, dist, N, z)
            p.state[:, 0] = z
        m.rand(eng, dist, N, z)              # x_t = a x_{t-1} + sx z
        p.state[:, 0] = p.state[:, 0] * a + z * sx
        d = (float(y[t]) - p.state[:, 0].to(torch.float64)) / sy
        p.weight().add_log(d * d * -0.5 + c)
        lnc.append(p.weight().log_nc())

    s.selection(sel)
    s.resample(m.SYSTEMATIC)
    s.resample_threshold(0.5)
    s.iterate(len(y))
    ll = sum(lnc)
    exact = kalman.loglik(y, a, sx, sy)
    print("kalman %.6f generic f32 %.6f" % (exact, ll))
    assert abs(ll - exact) < 0.3
    assert s.particle().state.dtype == torch.float32


def test_particle_select_and_errors(dev):
    import torch
    m = _m()
    p = m.Particle(10, 3, torch.float64)
    p.state.copy_(torch.arange(30, dtype=torch.float64).reshape(10, 3))
    idx = torch.tensor([9, 0, 0, 4], dtype=torch.int32, device="cuda")
    p.select(4, idx)
    assert p.size() == 4 and p.weight().size() == 4
    assert np.array_equal(p.state.cpu().numpy(),
                          np.arange(30.0).reshape(10, 3)[[9, 0, 0, 4]])
    assert np.allclose(p.weight().data.cpu().numpy(), 0.25)
    assert p.weight().ess() == 4.0
    with pytest.raises(m.MCKLError):
        p.select(4, idx.to(torch.int64))
    with pytest.raises(m.MCKLError):   # row 4 does not exist any more
        p.select(2, torch.tensor([0, 4], dtype=torch.int32, device="cuda"))
    with pytest.raises(m.MCKLError):
        p.select(1, torch.tensor([-1], dtype=torch.int32, device="cuda"))
    with pytest.raises(m.MCKLError):
        m.Particle(4, 1, torch.int32)
    s = m.GenericSMCSampler(8)
    with pytest.raises(m.MCKLError):
        s.resample(7)
    with pytest.raises(m.MCKLError):
        s.selection(3.5)
    assert s.resample_threshold() == -math.inf
    assert m.GenericSMCSampler.resample_threshold_always() == math.inf


def test_generic_pmmh_matches_oracle(dev):
    """GenericPMCMCMutation (PMCMCMutation over the generic sampler, the
    table-noise model as torch evaluation objects) against the oracle's
    restatement, which is pinned bit for bit to the reference's own
    PMCMCMutation: the same acceptance decisions, parameters to the Normal
    fill's rounding and log normalising constants to 1e-10."""
    import torch
    from mckl_amd import pmcmc
    m = _m()
    N, T, iters = 300, 12, 25
    scheme, thr = orc.SYSTEMATIC, 0.5
    y, z = _case(N, T, 40 + scheme)
    key = (777, scheme)

    def prior(q):
        return 0.0 if abs(q[0]) < 1 and q[1] > 0 and q[2] > 0 else -math.inf

    mut = m.GenericPMCMCMutation(N, T, prior, rng=m.Philox4x32(key=key))
    zd = torch.from_numpy(z).cuda()

    def sel(t, p):
        th = p.param
        x = p.state[:, 0]
        xn = (zd[:N].clone() if t == 0
              else x * float(th[0]) + zd[t * N:(t + 1) * N] * float(th[1]))
        p.state[:, 0] = xn
        d = (float(y[t]) - xn) / float(th[2])
        v = (d * d) * -0.5
        p.add_log_nc(torch.log((p.weight().data * torch.exp(v)).sum()))
        p.weight().add_log(v)

    mut.pf().selection(sel)
    mut.pf().resample(scheme)
    mut.pf().resample_threshold(thr)
    mut.update(pmcmc.random_walk((0.08, 0.08, 0.08)))
    param = np.array([0.7, 0.9, 0.6])
    acc, pars, lncs = [], [], []
    for it in range(iters):
        acc.append(mut(it, param))
        pars.append(param.copy())
        lncs.append(mut.pf().particle().log_nc())
    oacc, opar, olnc = orc.generic_pmmh_lg(
        N, T, iters, scheme, thr, key, (0.08, 0.08, 0.08), (0.7, 0.9, 0.6),
        y, z)
    print("accepted", sum(acc), "log_nc", lncs[-1], olnc[-1])
    assert acc == oacc.tolist()
    assert np.allclose(np.array(pars), opar, rtol=1e-13, atol=0)
    assert np.allclose(lncs, olnc, rtol=1e-10, atol=1e-10)
    assert 0 < sum(acc) < iters - 1
