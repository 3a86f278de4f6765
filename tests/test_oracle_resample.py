"""Pin the oracle's Weight and resampling: the compiled reference's fixtures,
the reference's own invariants (example/algorithm/include/
algorithm_resample_transform.hpp:37-80), and an exhaustive check of the
rep->index transform on every small replication vector.  CPU only."""
import itertools
import os

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cases():
    fx = np.load(os.path.join(GOLDEN, "resample_cases.npz"))
    k = 0
    while "c%d_meta" % k in fx:
        yield fx, "c%d" % k
        k += 1


def test_resample_fixtures():
    """ResampleAlgorithm + resample_trans_rep_index from the compiled
    reference, for the six schemes (the residual family through
    resample_trans_residual) at several (N, M)."""
    n = 0
    for fx, p in _cases():
        scheme, N, M, k0 = [int(v) for v in fx[p + "_meta"]]
        e = orc.engine(orc.PHILOX4X32, [k0, 0])
        rep, idx = orc.resample(scheme, e, fx[p + "_w"], M)
        assert np.array_equal(rep, fx[p + "_rep"]), p
        assert np.array_equal(idx, fx[p + "_idx"]), p
        e = orc.engine(orc.PHILOX4X32, [k0, 0])
        u = orc.u01_sequence(scheme, e, M)
        assert np.array_equal(u, fx[p + "_u"]), p
        n += 1
    assert n == 48


def check_rep_index(rep, idx, M):
    """Invariants that determine the transform uniquely: own slot kept when
    rep > 0, each source appears rep times, and extra copies are handed out
    in ascending source order to zero slots, then to slots >= N."""
    N = len(rep)
    K = min(N, M)
    assert len(idx) == M
    for d in range(K):
        if rep[d] > 0:
            assert idx[d] == d
    counts = np.bincount(np.asarray(idx, np.int64), minlength=N)
    assert np.array_equal(counts[:N], np.asarray(rep, np.int64))
    demand = [d for d in range(K) if rep[d] == 0] + list(range(K, M))
    vals = [int(idx[d]) for d in demand]
    assert vals == sorted(vals)


def test_rep_index_exhaustive_small():
    """Every replication vector with N <= 5 and M in {N-2..N+2}."""
    total = 0
    for N in range(1, 6):
        for M in range(max(1, N - 2), N + 3):
            for rep in itertools.product(range(M + 1), repeat=N):
                if sum(rep) != M:
                    continue
                idx = orc.rep_index(np.array(rep, np.uint64), M)
                check_rep_index(rep, idx, M)
                total += 1
    assert total > 1000


def test_rep_index_random_large():
    rng = np.random.default_rng(3)
    for N, M in ((1000, 1000), (4096, 5000), (5000, 4096), (70000, 70000)):
        w = rng.random(N) ** 6
        rep = rng.multinomial(M, w / w.sum()).astype(np.uint64)
        idx = orc.rep_index(rep, M)
        check_rep_index(rep, idx, M)


def test_trans_u01_rep_invariants():
    rng = np.random.default_rng(5)
    for N, M in ((1, 7), (2, 2), (1000, 1000), (300, 700), (700, 300)):
        w = rng.random(N)
        w /= w.sum()
        u = np.sort(rng.random(M))
        rep = orc.trans_u01_rep(w, u)
        assert int(rep.sum()) == M
        # rep_i counts uniforms in [S_{i-1}, S_i) for the sequential sums
        S = np.cumsum(w)
        for i in (0, N // 2, N - 2):
            if 0 <= i < N - 1:
                lo = S[i - 1] if i > 0 else 0.0
                assert rep[i] == np.sum((u >= lo) & (u < S[i]))


def test_weight_fixtures():
    """Weight::set_log / add_log from the compiled reference (bit-exact)."""
    fx = np.load(os.path.join(GOLDEN, "weight_cases.npz"))
    k = 0
    while "w%d_v" % k in fx:
        W, ess = orc.weight_set_log(fx["w%d_v" % k])
        assert np.array_equal(W, fx["w%d_W" % k])
        assert ess == fx["w%d_ess" % k][0]
        W2, ess2, lnc = orc.weight_add_log(W, fx["w%d_v2" % k])
        assert np.array_equal(W2, fx["w%d_W2" % k])
        assert ess2 == fx["w%d_ess2" % k][0]
        # the log normaliser increment is log sum W exp(v)
        ref = np.log(np.sum(W * np.exp(fx["w%d_v2" % k])))
        assert abs(lnc - ref) < 1e-12 * max(1.0, abs(ref))
        k += 1
    assert k == 7


def test_weight_closed_form():
    W, ess = orc.weight_set_log(np.zeros(4))
    assert np.array_equal(W, np.full(4, 0.25)) and ess == 4.0
    p = np.array([0.1, 0.2, 0.3, 0.4])
    W, ess = orc.weight_set_log(np.log(p))
    assert np.allclose(W, p, rtol=1e-15)
    assert abs(ess - 1.0 / np.sum(p * p)) < 1e-12


def test_exact_prefix_merge_against_rational_arithmetic():
    """The oracle's exactly-rounded-prefix merge (the device's definition of
    the cumulative weights): S_i = the exact rational sum of w_0..w_i
    rounded once to double, rep_i = #{j : S_{i-1} <= u_j < S_i}, last
    rep = M - j (resample.hpp:374-407).  Weights spanning 2^-60..1 make the
    sequential FP64 sums differ from the exact ones, and uniforms placed on
    and next to the exact boundaries exercise every comparison."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    differs = 0
    for N, M in ((2, 3), (17, 40), (200, 150), (300, 300)):
        w = rng.random(N) * np.exp2(-rng.integers(0, 60, N))
        w /= w.sum()
        acc = Fraction(0)
        S = []
        for x in w[:-1]:
            acc += Fraction(float(x))
            S.append(float(acc))
        pool = np.concatenate([rng.random(M), np.array(S),
                               np.nextafter(np.array(S), 0.0)])
        u = np.sort(rng.choice(pool, M, replace=False))
        want = np.zeros(N, np.uint64)
        j = 0
        for i, s in enumerate(S):
            while j < M and u[j] < s:
                want[i] += 1
                j += 1
        want[N - 1] = M - j
        got = orc.trans_u01_rep(w, u, exact_prefix=True)
        assert np.array_equal(got, want), (N, M)
        seq = orc.trans_u01_rep(w, u)
        assert int(seq.sum()) == M
        differs += int(not np.array_equal(seq, want))
    # the sequential sums do round differently here, so a merge on them
    # would fail the check above
    assert differs > 0
