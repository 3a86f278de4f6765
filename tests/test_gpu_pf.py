"""The reference's own particle-filter example
(example/algorithm/include/algorithm_pf.hpp:38-190, BackendSEQ) replayed on
the device (tests/cpp/algorithm_pf.hip: SMCSampler<double> with a ColMajor
4-column state, Normal fills on the range's engine, device functors for the
vmf arithmetic, Stratified resampling at 0.5) against the runs of the
reference itself in tests/golden/pf_reference.npz (tests/golden/make_pf.py).

Bit-exact: the engines' keys (SeedGenerator), the size history and every
resampling decision.  Within tolerance: the ESS history, the final state and
weights.  The device's Normal<double> values are the correctly rounded
Box-Muller values and the reference's carry glibc's log/sin/cos, which
misround about 3 in 10^4 of them by one ulp; those ulps ride along in the
states (sums of increments) and the log1p weights.  Measured on the MI355X:
99.7-100% of the final states bit-identical, the rest within 2.3e-16; ESS
within 1.4e-13 relative.  A single differing ancestor would put a state off
by O(1).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
PROG = os.path.join(ROOT, "tests", "cpp", "algorithm_pf")
FIX = np.load(os.path.join(GOLDEN, "pf_reference.npz"))
CASES = [(1000, 1), (10000, 7), (100000, 1)]


def test_fixture_keys_are_the_seed_generators():
    """CPU: the reference run's engine 0 and Particle::rng() keys are keys 0
    and N of the SeedGenerator sequence from internal seed `seed`, which the
    device program derives with mckl_hip_seed_keys."""
    for N, seed in CASES:
        k = orc.seed_keys(orc.PHILOX4X32, seed, N + 1)
        want = FIX["N%d_s%d_keys" % (N, seed)]
        assert k[0, :2].tolist() == want[:2].tolist()
        assert k[N, :2].tolist() == want[2:].tolist()


def test_fixture_shape():
    for N, seed in CASES:
        tag = "N%d_s%d" % (N, seed)
        assert len(FIX[tag + "_ess"]) == 100
        assert FIX[tag + "_state"].shape == (4, len(FIX[tag + "_rows"]))
        assert np.all(FIX[tag + "_ess"] > 0)


def run_device(N, seed, tmp_path):
    out = os.path.join(str(tmp_path), "pf_%d_%d.bin" % (N, seed))
    print("algorithm_pf N=%d seed=%d" % (N, seed), flush=True)
    r = subprocess.run([PROG, os.path.join(GOLDEN, "algorithm_pf.data"),
                        str(N), str(seed), "1000", out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = np.fromfile(out, dtype=np.float64)
    T = int(d[0])
    o = 1
    ess = d[o:o + T]
    o += T
    size = d[o:o + T]
    o += T
    state = d[o:o + 4 * N].reshape(4, N)
    o += 4 * N
    w = d[o:o + N]
    o += N
    keys = d[o:o + 4].view(np.uint64)
    return ess, size, state, w, keys


@pytest.mark.gpu
@pytest.mark.parametrize("N,seed", CASES)
def test_pf_matches_reference_run(N, seed, tmp_path):
    assert os.path.exists(PROG), "tests/cpp/algorithm_pf is not built"
    tag = "N%d_s%d" % (N, seed)
    ess, size, state, w, keys = run_device(N, seed, tmp_path)
    assert keys.tolist() == FIX[tag + "_keys"].tolist()
    assert size.astype(np.uint64).tolist() == FIX[tag + "_size"].tolist()
    ref_ess = FIX[tag + "_ess"]
    # the same resampling decisions (ESS < N/2) at every step
    assert np.array_equal(ess < 0.5 * N, ref_ess < 0.5 * N)
    assert np.allclose(ess, ref_ess, rtol=1e-11, atol=0)
    rows = FIX[tag + "_rows"]
    got = state[:, rows]
    want = FIX[tag + "_state"]
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13), \
        float(np.max(np.abs(got - want)))
    assert np.allclose(w[rows], FIX[tag + "_weight"], rtol=1e-12, atol=0)
    exact = float(np.mean(got == want))
    print("%s: states bit-identical %.4f, max |diff| %.3g, ESS max rel %.3g"
          % (tag, exact, float(np.max(np.abs(got - want))),
             float(np.max(np.abs(ess / ref_ess - 1)))), flush=True)
