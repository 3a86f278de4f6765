"""The fused samplers on models written by the caller
(mckl_amd/include/mckl_hip_fused.hpp; the models are in
tests/cpp/fused_model_test.hip):

* a caller-written linear-Gaussian model gives the built-in sampler's bits,
  plain and batched, for every scheme: the model enters the kernels through
  the model type alone;
* the nonstationary growth model (a model the library does not have) on
  the plain fused sampler agrees with the oracle's restatement (model 3):
  the same resampling decisions, ESS to 1e-12 and log-likelihood to 1e-10,
  as the built-in models' parity tests ask;
* the batched form of the growth model gives each filter the plain run's
  bits;
* the plain sampler's moment estimates on the growth model agree with the
  oracle's;
* a model with a two-float state (the oracle's model 4, pinned by a
  Kalman filter) gives the oracle's runs, plain and batched;
* the reference's own particle-filter model (a four-float state) agrees
  with the reference's run statistically.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "libfused_model_test.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")
SCHEMES = [0, 1, 2, 3, 4, 5]
LG = (0.9, 1.0, 0.5)
GROWTH = (np.sqrt(10.0), 1.0, np.sqrt(5.0))  # (s, sy, sd0)


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(scope="module")
def fl(dev):
    import mckl_amd  # noqa: F401  (loads the library and the HIP runtime)
    lib = ctypes.CDLL(LIB)
    lib.fused_last_error.restype = ctypes.c_char_p
    return lib


def growth_data(T, seed):
    """Observations of the growth model and its drifts u_t = 8 cos(1.2 t),
    concatenated (the oracle's model 3 layout)."""
    rng = np.random.default_rng(seed)
    s, sy, sd0 = GROWTH
    x = sd0 * rng.standard_normal()
    y = np.zeros(T)
    u = 8.0 * np.cos(1.2 * np.arange(1, T + 1))
    for t in range(T):
        x = 0.5 * x + 25.0 * x / (1.0 + x * x) + u[t] + s * rng.standard_normal()
        y[t] = x * x / 20.0 + sy * rng.standard_normal()
    return np.concatenate([y, u])


def plain(fl, model, th, y, T, N, seed, thr, scheme):
    th = np.ascontiguousarray(th, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    ll = ctypes.c_double()
    ess = np.zeros(T)
    res = np.zeros(T, np.int32)
    rc = fl.fused_plain_run(model, P(th), P(y), ctypes.c_size_t(T),
                            ctypes.c_size_t(N), ctypes.c_uint64(seed),
                            ctypes.c_double(thr), scheme, ctypes.byref(ll),
                            P(ess), P(res))
    assert rc == 0, fl.fused_last_error()
    return ll.value, ess, res


def batch(fl, model, th, y, T, N, seeds, thr, scheme):
    nf = len(N)
    th = np.ascontiguousarray(th, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    Nv = np.ascontiguousarray(N, np.uint32)
    sv = np.ascontiguousarray(seeds, np.uint64)
    ll = np.zeros(nf)
    rc = fl.fused_batch_run(model, ctypes.c_size_t(nf), P(Nv), P(th), P(y),
                            ctypes.c_size_t(T), P(sv), ctypes.c_double(thr),
                            scheme, P(ll))
    assert rc == 0, fl.fused_last_error()
    return ll


@pytest.mark.parametrize("scheme", SCHEMES)
def test_user_lg_equals_builtin(fl, scheme):
    import mckl_amd as m
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"][:40]
    for N, seed in ((4096, 3), (5000, 4), (1 << 16, 5)):
        ll, ess, res = plain(fl, 1, LG, y, len(y), N, seed, 0.5, scheme)
        s = m.SMCSampler(m.MODEL_LG, LG, y, N, seed, 0.5, scheme)
        s.iterate()
        bl, bess, bres = s.results()
        print(scheme, N, ll, bl, int(res.sum()))
        assert ll == bl
        assert np.array_equal(ess, bess) and np.array_equal(res, bres)


@pytest.mark.parametrize("scheme", [1, 2, 0, 5])
def test_user_lg_batch_equals_builtin(fl, scheme):
    import mckl_amd as m
    y = np.load(os.path.join(GOLDEN, "observations.npz"))["lg_y"][:30]
    T = len(y)
    N = [300, 2048, 2049, 5000, 17000]
    seeds = [11, 12, 13, 14, 15]
    th = np.array([LG, (0.8, 1.1, 0.6), LG, (0.95, 0.7, 0.4), LG])
    ll = batch(fl, 1, th, np.tile(y, len(N)), T, N, seeds, 0.5, scheme)
    b = m.SMCBatch(m.MODEL_LG, N, th, np.tile(y, (len(N), 1)), seeds, 0.5,
                   scheme)
    b.run()
    want = b.log_likelihood()
    print(ll, want)
    assert np.array_equal(ll, want)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_growth_model_vs_oracle(fl, scheme):
    T = 50
    y = growth_data(T, 21)
    for N, seed in ((4096, 7), (3000, 8)):
        ll, ess, res = plain(fl, 3, GROWTH, y, T, N, seed, 0.5, scheme)
        oll, oess, ores = orc.smc_run(3, GROWTH, y, N, seed, 0.5, scheme,
                                      nthreads=8, history=True)
        print(scheme, N, ll, oll, int(res.sum()))
        assert 0 < res.sum() <= T
        assert np.array_equal(res, ores)
        assert np.allclose(ess, oess, rtol=1e-12, atol=0)
        assert abs(ll - oll) < 1e-10 * max(1.0, abs(oll))


CV2 = (0.3, 0.2, 0.8)  # (sp, sv, sy)


def cv2_data(T, seed):
    """Observations of the two-component constant-velocity model."""
    rng = np.random.default_rng(seed)
    sp, sv, sy = CV2
    x = rng.standard_normal(2)
    y = np.zeros(T)
    for t in range(T):
        x = np.array([x[0] + x[1], x[1]]) + np.array([sp, sv]) * \
            rng.standard_normal(2)
        y[t] = x[0] + sy * rng.standard_normal()
    return y


@pytest.mark.parametrize("scheme", SCHEMES)
def test_vector_state_vs_oracle(fl, scheme):
    """A model with a two-float state (dim = 2: states by component, the
    normals i D + d) gives the oracle's model 4 run: same decisions, ESS to
    1e-12, log Z to 1e-10."""
    T = 40
    y = cv2_data(T, 31)
    for N, seed in ((4096, 7), (3000, 8), (2049, 9)):
        ll, ess, res = plain(fl, 4, CV2, y, T, N, seed, 0.5, scheme)
        oll, oess, ores = orc.smc_run(4, CV2, y, N, seed, 0.5, scheme,
                                      nthreads=8, history=True)
        print(scheme, N, ll, oll, int(res.sum()))
        assert 0 < res.sum() <= T
        assert np.array_equal(res, ores)
        assert np.allclose(ess, oess, rtol=1e-12, atol=0)
        assert abs(ll - oll) < 1e-10 * max(1.0, abs(oll))


def test_vector_state_batch_equals_plain(fl):
    T = 30
    N = [256, 2048, 4097, 9000, 640]
    seeds = [41, 42, 43, 44, 45]
    ys = [cv2_data(T, 50 + f) for f in range(len(N))]
    th = np.array([CV2] * len(N))
    th[1, 1] = 0.4
    th[3, 2] = 1.2
    for scheme in (1, 2, 5):
        ll = batch(fl, 4, th, np.concatenate(ys), T, N, seeds, 0.5, scheme)
        for f in range(len(N)):
            lp, _, _ = plain(fl, 4, th[f], ys[f], T, N[f], seeds[f], 0.5,
                             scheme)
            assert ll[f] == lp, (scheme, f, ll[f], lp)


def test_vector_state_moments_vs_oracle(fl):
    """The moment estimates of a vector state are those of its first
    component (the oracle's too)."""
    T, k = 20, 2
    y = cv2_data(T, 33)
    th = np.ascontiguousarray(CV2, np.float64)
    yy = np.ascontiguousarray(y, np.float64)
    for N in (1, 3000):
        ll = ctypes.c_double()
        est = np.zeros(T * k)
        rc = fl.fused_plain_moments(4, P(th), P(yy), ctypes.c_size_t(T),
                                    ctypes.c_size_t(N), ctypes.c_uint64(9),
                                    ctypes.c_double(0.5), 2, k,
                                    ctypes.byref(ll), P(est))
        assert rc == 0, fl.fused_last_error()
        oll, _, _, oest = orc.smc_run(4, CV2, y, N, 9, 0.5, 2, nthreads=8,
                                      moments=k)
        est = est.reshape(T, k)
        assert abs(ll.value - oll) < 1e-10 * max(1.0, abs(oll))
        assert np.all(np.abs(est - oest) < 1e-9 * (np.abs(oest) + 1.0))


def test_reference_pf_model(fl):
    """The reference's particle-filter example model (four-float state,
    algorithm_pf.hpp:91-150) on the fused sampler, against the reference's
    own run (pf_reference.npz, N = 10^4, BackendSEQ): the last step's
    posterior mean of the position p from 2^20 particles lies within 4
    standard deviations of the fused estimate at N = 10^4 (measured over 16
    seeds) of the reference's weighted mean.  Same model, other noise and
    FP32 states: a statistical check."""
    xy = np.loadtxt(os.path.join(GOLDEN, "algorithm_pf.data")).ravel()
    T = len(xy) // 2
    fix = np.load(os.path.join(GOLDEN, "pf_reference.npz"))
    p, w = fix["N10000_s7_state"][0], fix["N10000_s7_weight"]
    ref = float(np.sum(w * p) / np.sum(w))

    def run(N, seed):
        ll = ctypes.c_double()
        ess = np.zeros(T)
        res = np.zeros(T, np.int32)
        post = np.zeros(2)
        ms = ctypes.c_double()
        xx = np.ascontiguousarray(xy, np.float64)
        rc = fl.fused_pf_run(P(xx), ctypes.c_size_t(T), ctypes.c_size_t(N),
                             ctypes.c_uint64(seed), ctypes.byref(ll), P(ess),
                             P(res), P(post), ctypes.byref(ms))
        assert rc == 0, fl.fused_last_error()
        return ll.value, ess, res, post, ms.value

    small = np.array([run(10000, 100 + s)[3][0] for s in range(16)])
    sd = small.std(ddof=1)
    ll, ess, res, post, ms = run(1 << 20, 5)
    print("p mean: fused 2^20 %.6f (%.3f ms per run, %d resampling steps), "
          "reference N=1e4 %.6f, sd at 1e4 %.2e" % (post[0], ms,
                                                   int(res.sum()), ref, sd))
    assert np.all(ess > 0) and 0 < res.sum() <= T
    assert post[1] >= post[0] ** 2 * (1 - 1e-6)
    assert abs(post[0] - ref) < 4 * sd
    assert abs(small.mean() - post[0]) < 4 * sd / np.sqrt(len(small)) + \
        1e-3 * abs(post[0])


@pytest.mark.parametrize("scheme", [1, 2, 0])
def test_growth_moments_vs_oracle(fl, scheme):
    """FusedSMCSampler::estimate_moments on the caller's growth model: the
    selection estimates of x, x^2, x^3 at every step agree with the oracle's
    (same particles and weights; the oracle sums W_i x_i^j in order, the
    device in double-double)."""
    T, k = 30, 3
    y = growth_data(T, 23)
    for N in (1, 3000, 4096):
        th = np.ascontiguousarray(GROWTH, np.float64)
        yy = np.ascontiguousarray(y, np.float64)
        ll = ctypes.c_double()
        est = np.zeros(T * k)
        rc = fl.fused_plain_moments(3, P(th), P(yy), ctypes.c_size_t(T),
                                    ctypes.c_size_t(N), ctypes.c_uint64(9),
                                    ctypes.c_double(0.5), scheme, k,
                                    ctypes.byref(ll), P(est))
        assert rc == 0, fl.fused_last_error()
        oll, _, _, oest = orc.smc_run(3, GROWTH, y, N, 9, 0.5, scheme,
                                      nthreads=8, moments=k)
        est = est.reshape(T, k)
        assert abs(ll.value - oll) < 1e-10 * max(1.0, abs(oll))
        scale = np.abs(oest) + 1.0
        assert np.all(np.abs(est - oest) < 1e-9 * scale), (N, scheme)
        if N == 1:
            assert np.allclose(est[:, 1], est[:, 0] ** 2, rtol=1e-6)


def test_growth_batch_equals_plain(fl):
    T = 40
    N = [256, 2048, 4097, 9000, 640]
    seeds = [31, 32, 33, 34, 35]
    ys = [growth_data(T, 40 + f) for f in range(len(N))]
    th = np.array([GROWTH] * len(N))
    th[1, 0] = 2.5
    th[3, 1] = 1.5
    for scheme in (1, 2, 4):
        ll = batch(fl, 3, th, np.concatenate(ys), T, N, seeds, 0.5, scheme)
        for f in range(len(N)):
            lp, _, _ = plain(fl, 3, th[f], ys[f], T, N[f], seeds[f], 0.5,
                             scheme)
            assert ll[f] == lp, (scheme, f, ll[f], lp)
