"""GPU Weight and resampling primitives against the compiled-reference
fixtures and the oracle.

Replication counts come from FP64 prefix sums.  The device's cumulative
weights are the exact prefix sums rounded once (a double-double prefix),
the reference's are sequential FP64 sums.  Parity is therefore checked two
ways:
  * bit-exact: rep against the oracle's merge on the exactly rounded prefix
    (the device's definition) of the same weights and the device's own
    uniforms, and idx against the oracle's resample_trans_rep_index of it;
  * against the reference's own rep (fixtures, or the oracle's sequential
    form): every difference must be a uniform within the sequential sum's
    error (N eps) of an exact cumulative weight, and the cases within 1e-15
    relative and beyond it are counted and reported.
"""
import math
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _m():
    import mckl_amd as m
    return m


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def boundary_report(w, us, rep_ref, rep_gpu, tol=None):
    """Differences between the reference's rep and the device's: each must
    have a uniform (of any sequence in `us`) within `tol` (default N eps)
    relative of an exact cumulative weight at that particle.  Returns
    (differences, of them within 1e-15 relative, largest relative
    distance)."""
    diff = np.nonzero(rep_ref.astype(np.int64) != rep_gpu.astype(np.int64))[0]
    if len(diff) == 0:
        return 0, 0, 0.0
    exact = np.cumsum(np.asarray(w, np.longdouble))
    N = len(w)
    tol = N * np.finfo(np.float64).eps if tol is None else tol
    within = 0
    worst = 0.0
    for i in diff:
        d = 1.0
        for u in us:
            S = float(exact[min(i, N - 1)])
            Sp = float(exact[i - 1]) if i > 0 else 0.0
            d = min(d, np.min(np.abs(u - S)) / S if S else 1.0,
                    np.min(np.abs(u - Sp)) / Sp if Sp else 1.0)
        assert d <= tol, (i, d)
        within += d <= 1e-15
        worst = max(worst, d)
    return len(diff), within, worst


def strict_rep_idx(w, u, M, rep_gpu, idx_gpu):
    """rep bit-exact against the exactly rounded prefix merge, idx against
    the stable rep -> index of it."""
    want = orc.trans_u01_rep(w, u, len(w), M, exact_prefix=True)
    assert np.array_equal(rep_gpu.astype(np.uint64), want)
    if idx_gpu is not None:
        assert np.array_equal(idx_gpu.astype(np.uint64),
                              orc.rep_index(want, M))


def test_resample_fixtures(dev):
    import torch
    m = _m()
    fx = np.load(os.path.join(GOLDEN, "resample_cases.npz"))
    k = 0
    tot = np.zeros(3, np.int64)
    while "c%d_meta" % k in fx:
        p = "c%d" % k
        scheme, N, M, k0 = [int(v) for v in fx[p + "_meta"]]
        w = _t(fx[p + "_w"], dev)
        rep = torch.zeros(N, dtype=torch.int32, device=dev)
        idx = torch.zeros(M, dtype=torch.int32, device=dev)
        rng = m.Philox4x32(key=[k0, 0])
        alg = {0: m.ResampleMultinomial, 1: m.ResampleStratified,
               2: m.ResampleSystematic, 3: m.ResampleResidual,
               4: m.ResampleResidualStratified,
               5: m.ResampleResidualSystematic}[scheme]()
        alg(N, M, rng, w, rep, idx)
        # uniform sequence itself
        rng2 = m.Philox4x32(key=[k0, 0])
        u = torch.zeros(max(M, 1), dtype=torch.float64, device=dev)
        m.u01_rand_sequence(scheme % 3, rng2, M, u)
        ug = u.cpu().numpy()[:M]
        if scheme % 3 == 0:
            # sorted uniforms: prefix of log(u)/(N-j), rounding may differ
            assert np.allclose(ug, fx[p + "_u"], rtol=1e-12, atol=1e-15)
        else:
            assert np.array_equal(ug, fx[p + "_u"])
        rg = rep.cpu().numpy().view(np.uint32).astype(np.uint64)
        ig = idx.cpu().numpy().view(np.uint32).astype(np.uint64)
        if scheme >= 3:
            # residual family: integral parts exact, R uniforms over the
            # normalised residuals
            assert np.array_equal(rg, fx[p + "_rep"]), p
            assert np.array_equal(ig, fx[p + "_idx"]), p
            k += 1
            continue
        strict_rep_idx(fx[p + "_w"], ug, M, rg, ig)
        tol = None if scheme else N * np.finfo(np.float64).eps + 1e-12
        tot += boundary_report(fx[p + "_w"], (fx[p + "_u"], ug),
                               fx[p + "_rep"], rg, tol)[:2] + (0,)
        k += 1
    print("replication differences from the reference (all at cumulative-"
          "weight boundaries): %d, of them within 1e-15 relative: %d" %
          (tot[0], tot[1]))


@pytest.mark.parametrize("N,M", [(1, 1), (1, 5), (5, 1), (2, 2), (1000, 1000),
                                 (4097, 5000), (5000, 4097), (70000, 70000)])
def test_rep_index_vs_oracle(dev, N, M):
    import torch
    m = _m()
    rng = np.random.default_rng(N * 31 + M)
    w = rng.random(N) ** 5
    rep = rng.multinomial(M, w / w.sum()).astype(np.uint32)
    idx = torch.zeros(M, dtype=torch.int32, device=dev)
    m.resample_trans_rep_index(N, M, _t(rep.view(np.int32), dev), idx)
    want = orc.rep_index(rep.astype(np.uint64), M)
    assert np.array_equal(idx.cpu().numpy().view(np.uint32).astype(np.uint64),
                          want)


def test_trans_u01_rep_vs_oracle(dev):
    import torch
    m = _m()
    rng = np.random.default_rng(11)
    for N, M in ((2, 2), (1000, 1000), (3000, 7000), (7000, 3000), (1, 9)):
        w = rng.random(N)
        w /= w.sum()
        u = np.sort(rng.random(M))
        rep = torch.zeros(N, dtype=torch.int32, device=dev)
        m.resample_trans_u01_rep(N, M, _t(w, dev), _t(u, dev), rep)
        rg = rep.cpu().numpy().astype(np.uint64)
        strict_rep_idx(w, u, M, rg, None)
        ro = orc.trans_u01_rep(w, u)
        boundary_report(w, (u,), ro, rg)
        assert int(rg.sum()) == M


def test_weight_fixtures(dev):
    m = _m()
    fx = np.load(os.path.join(GOLDEN, "weight_cases.npz"))
    k = 0
    while "w%d_v" % k in fx:
        v = fx["w%d_v" % k]
        n = len(v)
        W = m.Weight(n)
        W.set_log(_t(v, dev))
        got = W.data.cpu().numpy()
        assert np.allclose(got, fx["w%d_W" % k], rtol=1e-13, atol=0)
        assert abs(W.ess() - fx["w%d_ess" % k][0]) <= 1e-12 * n
        W.add_log(_t(fx["w%d_v2" % k], dev))
        got2 = W.data.cpu().numpy()
        assert np.allclose(got2, fx["w%d_W2" % k], rtol=1e-12, atol=0)
        assert abs(W.ess() - fx["w%d_ess2" % k][0]) <= 1e-11 * n
        _, _, lnc = orc.weight_add_log(fx["w%d_W" % k], fx["w%d_v2" % k])
        assert abs(W.log_nc() - lnc) < 1e-12 * max(1.0, abs(lnc))
        k += 1


def test_select(dev):
    import torch
    m = _m()
    src = torch.arange(3000, dtype=torch.float32, device=dev)
    idx = torch.tensor(np.random.default_rng(1).integers(0, 1000, 777),
                       dtype=torch.int32, device=dev)
    dst = torch.empty(777 * 3, dtype=torch.float32, device=dev)
    m.select(idx, src, dst, dim=3)
    i = idx.cpu().numpy()
    want = np.stack([3 * i, 3 * i + 1, 3 * i + 2], 1).reshape(-1)
    assert np.array_equal(dst.cpu().numpy(), want.astype(np.float32))


@pytest.mark.parametrize("scheme", range(6))
@pytest.mark.parametrize("N,M", [((1 << 20) + 12345, (1 << 20) + 12345),
                                 (600001, 900007), (900007, 600001),
                                 ((1 << 22) + 777, (1 << 22) + 777),
                                 ((1 << 22) + 5, 3000001)])
def test_resample_large_vs_oracle(dev, scheme, N, M):
    """Sizes past 256 tiles of 2048, where every workgroup's cross-tile
    loops (tile offsets of the weights, the rep -> index offsets, the zero
    total) take several tiles per thread, and past 1024 tiles, where one
    workgroup writes every tile's offsets for the others to read: rep and
    idx against the oracle (the residual family exactly; the others up to
    replication changes at a uniform within rounding of a cumulative
    weight), and idx a stable realisation of rep."""
    import torch
    m = _m()
    rng = np.random.default_rng(N + 7 * scheme)
    w = rng.random(N) ** 4
    w /= w.sum()
    key = [9000 + scheme, 0]
    rep = torch.zeros(N, dtype=torch.int32, device=dev)
    idx = torch.zeros(M, dtype=torch.int32, device=dev)
    alg = {0: m.ResampleMultinomial, 1: m.ResampleStratified,
           2: m.ResampleSystematic, 3: m.ResampleResidual,
           4: m.ResampleResidualStratified,
           5: m.ResampleResidualSystematic}[scheme]()
    alg(N, M, m.Philox4x32(key=key), _t(w, dev), rep, idx)
    rg = rep.cpu().numpy().view(np.uint32).astype(np.uint64)
    ig = idx.cpu().numpy().view(np.uint32).astype(np.uint64)
    e = orc.engine(orc.PHILOX4X32, key)
    orep, oidx = orc.resample(scheme, e, w, M)
    assert rg.sum() == M
    # idx realises rep: slot d < min(N, M) keeps d when rep_d > 0
    assert np.array_equal(np.bincount(ig.astype(np.int64), minlength=N), rg)
    K = min(N, M)
    keep = rg[:K] > 0
    assert np.array_equal(ig[:K][keep], np.arange(K)[keep])
    if scheme >= 3:
        assert np.array_equal(rg, orep)
        assert np.array_equal(ig, oidx)
        return
    e2 = orc.engine(orc.PHILOX4X32, key)
    u = orc.u01_sequence(scheme, e2, M)
    ud = torch.zeros(M, dtype=torch.float64, device=dev)
    m.u01_rand_sequence(scheme, m.Philox4x32(key=key), M, ud)
    ud = ud.cpu().numpy()
    strict_rep_idx(w, ud, M, rg, ig)
    tol = None if scheme else N * np.finfo(np.float64).eps + 1e-12
    nd, n15, worst = boundary_report(w, (u, ud), orep, rg, tol)
    print("scheme %d N=%d M=%d differences from the reference %d (within "
          "1e-15: %d, largest distance %.3g)" % (scheme, N, M, nd, n15, worst))


@pytest.mark.parametrize("N", [(1 << 20) + 3, 3 * (1 << 19) + 1])
def test_weight_large_vs_oracle(dev, N):
    """Weight::set_log / add_log past 256 tiles (the block maxima and the
    tile sums taken in every workgroup over several tiles per thread)
    against the oracle, at the fixture test's tolerances."""
    m = _m()
    rng = np.random.default_rng(N)
    v = rng.normal(size=N) * 3.0
    v2 = rng.normal(size=N) * 2.0
    W = m.Weight(N)
    W.set_log(_t(v, dev))
    w1, ess1 = orc.weight_set_log(v)
    assert np.allclose(W.data.cpu().numpy(), w1, rtol=1e-13, atol=0)
    assert abs(W.ess() - ess1) <= 1e-12 * N
    W.add_log(_t(v2, dev))
    w2, ess2, lnc = orc.weight_add_log(w1, v2)
    assert np.allclose(W.data.cpu().numpy(), w2, rtol=1e-12, atol=0)
    assert abs(W.ess() - ess2) <= 1e-11 * N
    assert abs(W.log_nc() - lnc) < 1e-12 * max(1.0, abs(lnc))


@pytest.mark.parametrize("scheme", range(6))
def test_resample_tiny_sizes(dev, scheme):
    """N = 1 (every copy of the one particle), M = 1 and N = M = 2 through
    mckl_hip_resample against the oracle: rep sums to M and idx realises
    it."""
    import torch
    m = _m()
    alg = {0: m.ResampleMultinomial, 1: m.ResampleStratified,
           2: m.ResampleSystematic, 3: m.ResampleResidual,
           4: m.ResampleResidualStratified,
           5: m.ResampleResidualSystematic}[scheme]()
    for N, M in ((1, 5), (5, 1), (2, 2), (1, 1)):
        w = np.arange(1, N + 1, dtype=np.float64)
        w /= w.sum()
        rep = torch.zeros(N, dtype=torch.int32, device=dev)
        idx = torch.zeros(M, dtype=torch.int32, device=dev)
        key = [31 + N, M]
        alg(N, M, m.Philox4x32(key=key), _t(w, dev), rep, idx)
        orep, oidx = orc.resample(scheme, orc.engine(orc.PHILOX4X32, key), w,
                                  M)
        rg = rep.cpu().numpy().astype(np.uint64)
        ig = idx.cpu().numpy().astype(np.uint64)
        assert rg.sum() == M and np.array_equal(rg, orep), (N, M, rg, orep)
        assert np.array_equal(ig, oidx), (N, M, ig, oidx)
