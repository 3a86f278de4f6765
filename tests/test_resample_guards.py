"""The resampling kernels' index arithmetic (mckl_amd/csrc/smc_index.hpp),
run on the CPU over adversarial counts (tests/cpp/rs_guard_host.cpp emulates
the tile stages around it and checks every array access).

* With the boundary counts the kernels compute (tile k's upper boundary is
  tile k+1's lower one), any raw counts -- non-monotone, all mass on one
  particle, off by one at tile ends, N = 1, 2 and tile boundaries -- become
  a replication vector (non-negative, summing to N), the parent indices
  equal the reference's stable rep -> index transform of it (the oracle's
  resample_trans_rep_index), and no access leaves its array.
* With boundary counts that disagree between neighbouring tiles (which the
  kernels never produce), no access leaves its array either.
* The plain sampler's form (zero slots coded by rank, resolved through a
  rank -> source table written at each tile's look-back offset) gives the
  same parents and stays inside its arrays on the same inputs.
* The guarded counts (a cumulative weight known in double to within a
  margin, counted from the uniforms' closed form) equal brute-force counts
  of the systematic and stratified uniforms at every weight they claim,
  with weights placed on, beside and between uniforms.
* The form before the clamp and the rank bound, replayed on the CPU, reads
  the zero-slot array past its end on such inputs: the out-of-range access
  behind the GPU fault of an earlier resampling-kernel variant.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "rs_guard_host.cpp")
LIB = os.path.join(ROOT, "tests", "cpp", "librs_guard_host.so")
HDR = os.path.join(ROOT, "mckl_amd", "csrc", "smc_index.hpp")
TILE = 2048


@pytest.fixture(scope="module")
def rg():
    if (not os.path.exists(LIB) or
            os.path.getmtime(LIB) < max(os.path.getmtime(SRC),
                                        os.path.getmtime(HDR))):
        subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off",
                        "-fPIC", "-shared", "-Wall", "-o", LIB, SRC],
                       check=True)
    lib = ctypes.CDLL(LIB)
    lib.rs_guard_run.restype = ctypes.c_uint64
    lib.rs_guard_src_run.restype = ctypes.c_uint64
    lib.guard_count_check.restype = ctypes.c_uint64
    return lib


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def run(rg, raw, clo, chi, guarded=True):
    N = len(raw)
    raw = np.ascontiguousarray(raw, np.uint32)
    clo = np.ascontiguousarray(clo, np.uint32)
    chi = np.ascontiguousarray(chi, np.uint32)
    idx = np.zeros(max(N, 1), np.uint32)
    rep = np.zeros(max(N, 1), np.uint32)
    bad = rg.rs_guard_run(ctypes.c_uint64(N), P(raw), P(clo), P(chi),
                          int(guarded), P(idx), P(rep))
    return int(bad), idx[:N], rep[:N]


def boundaries(raw, N, jitter=None, rng=None):
    """Boundary counts as the kernels take them: non-decreasing over tiles
    (counts at a prefix of non-negative weights) and chi[k] = clo[k+1]; with
    jitter, tile k+1's lower boundary differs from tile k's upper one."""
    nt = (N + TILE - 1) // TILE
    b = np.zeros(nt + 1, np.int64)
    for k in range(1, nt):
        b[k] = min(N, max(b[k - 1], raw[k * TILE - 1]))
    b[nt] = N
    clo = b[:nt].copy()
    chi = b[1:].copy()
    if jitter:
        d = rng.integers(-jitter, jitter + 1, nt)
        clo = np.clip(clo + d, 0, N)
        clo[0] = 0
    return clo, chi


def patterns(N, rng):
    sorted_counts = np.sort(rng.integers(0, N + 1, N))
    sorted_counts[-1] = N
    yield "monotone", sorted_counts
    yield "random", rng.integers(0, N + 1, N)
    yield "all_on_first", np.full(N, N)
    yield "all_on_last", np.concatenate([np.zeros(N - 1, np.int64), [N]])
    edge = sorted_counts.copy()
    for k in range(TILE - 1, N, TILE):  # off by one around tile ends
        for j in (k - 1, k, k + 1):
            if 0 <= j < N:
                edge[j] = min(N, max(0, edge[j] + rng.integers(-1, 2)))
    yield "tile_edges", edge
    yield "descending", np.arange(N, 0, -1)


SIZES = [1, 2, 3, 2047, 2048, 2049, 4101, 3 * 2048]


@pytest.mark.parametrize("N", SIZES)
def test_guarded_counts_make_a_replication_vector(rg, N):
    rng = np.random.default_rng(N)
    for name, raw in patterns(N, rng):
        clo, chi = boundaries(raw, N)
        bad, idx, rep = run(rg, raw, clo, chi)
        assert bad == 0, name
        assert rep.astype(np.int64).sum() == N, name
        assert np.all(rep <= N), name  # no wrapped negative numbers
        assert np.all(idx < N), name
        want = orc.rep_index(rep.astype(np.uint64), N)
        assert np.array_equal(idx.astype(np.uint64), want), name


@pytest.mark.parametrize("N", [2049, 4101, 3 * 2048, 5 * 2048 + 17])
def test_guarded_accesses_stay_inside_with_inconsistent_boundaries(rg, N):
    rng = np.random.default_rng(7 * N)
    for jitter in (1, 5, N):
        for name, raw in patterns(N, rng):
            clo, chi = boundaries(raw, N, jitter, rng)
            bad, idx, _ = run(rg, raw, clo, chi)
            assert bad == 0, (name, jitter)
            assert np.all(idx < N), (name, jitter)


def test_unguarded_form_reads_past_the_zero_slots(rg):
    """The pre-guard arithmetic replayed on the CPU.  Particle 2046's own
    prefix rounds above the prefix at the end of tile 0, so its count
    (2047) exceeds the count the tile's last particle takes from the
    boundary (2046): that particle's replication number is -1, the u32
    extra-copy prefix wraps to about 2^32 at the end of the tile, and the
    tile's rank loop reads the zero-slot array past its end -- on the GPU,
    an illegal address.  The guarded form confines the counts to [clo, chi]
    and stays inside every array."""
    N = 3 * 2048
    raw = np.arange(1, N + 1)      # one copy each ...
    clo = np.array([0, 2046, 4096])
    chi = np.array([2046, 4096, N])  # ... but tile 0 ends at count 2046
    bad_old, _, rep_old = run(rg, raw, clo, chi, guarded=False)
    assert bad_old > 0
    assert rep_old[2047] == 2 ** 32 - 1  # the wrapped -1
    bad, idx, rep = run(rg, raw, clo, chi)
    assert bad == 0
    assert rep.astype(np.int64).sum() == N and rep.max() <= N
    assert np.array_equal(idx.astype(np.uint64),
                          orc.rep_index(rep.astype(np.uint64), N))


def run_src(rg, raw, clo, chi):
    N = len(raw)
    raw = np.ascontiguousarray(raw, np.uint32)
    clo = np.ascontiguousarray(clo, np.uint32)
    chi = np.ascontiguousarray(chi, np.uint32)
    idx = np.zeros(max(N, 1), np.uint32)
    rep = np.zeros(max(N, 1), np.uint32)
    bad = rg.rs_guard_src_run(ctypes.c_uint64(N), P(raw), P(clo), P(chi),
                              P(idx), P(rep))
    return int(bad), idx[:N], rep[:N]


@pytest.mark.parametrize("N", SIZES + [5 * 2048 + 17])
def test_rank_source_form_equals_the_reference_transform(rg, N):
    rng = np.random.default_rng(3 * N + 1)
    for name, raw in patterns(N, rng):
        for jitter in (None, 1, N):
            clo, chi = boundaries(raw, N, jitter, rng)
            bad, idx, rep = run_src(rg, raw, clo, chi)
            assert bad == 0, (name, jitter)
            assert np.all(idx < N), (name, jitter)
            if jitter is None:
                assert rep.astype(np.int64).sum() == N, name
                want = orc.rep_index(rep.astype(np.uint64), N)
                assert np.array_equal(idx.astype(np.uint64), want), name
                _, idx_old, rep_old = run(rg, raw, clo, chi)
                assert np.array_equal(rep, rep_old), name
                assert np.array_equal(idx, idx_old), name


MARGIN = 64 * 2.0 ** -53  # kCountMargin


def weights_near_uniforms(u, rng, n):
    """Cumulative weights on uniforms, a few ulps either side, at the
    margin's edges, between uniforms and past both ends."""
    j = rng.integers(0, len(u), n)
    base = u[j]
    k = rng.integers(-300, 301, n).astype(np.float64)
    out = [base * (1.0 + k * 2.0 ** -53),
           base * (1.0 + MARGIN * rng.uniform(-3, 3, n)),
           rng.uniform(0.0, 1.0, n),
           np.array([0.0, 1e-300, u[0], u[-1], 1.0, 1.5])]
    return np.ascontiguousarray(np.concatenate(out))


@pytest.mark.parametrize("stratified", [0, 1])
@pytest.mark.parametrize("M", [1, 2, 3, 7, 1000, 2048, 4101, 1 << 20,
                               (1 << 20) + 12345])
def test_guarded_counts_equal_brute_force(rg, stratified, M):
    rng = np.random.default_rng(M * 2 + stratified)
    u01 = float(rng.uniform())
    us = np.ascontiguousarray(rng.uniform(0.0, 1.0, M))
    delta = 1.0 / M
    if stratified:
        u = delta * (us + np.arange(M, dtype=np.float64))
    else:
        u = np.arange(M, dtype=np.float64) * delta + u01 * delta
    S = weights_near_uniforms(u, rng, 20000)
    nok = ctypes.c_uint64(0)
    bad = rg.guard_count_check(stratified, ctypes.c_uint32(M),
                               ctypes.c_double(u01), P(us), P(S),
                               ctypes.c_uint64(len(S)),
                               ctypes.c_double(MARGIN), ctypes.byref(nok))
    assert bad == 0
    # the guard decides most weights itself (those on or within a few ulps
    # of a uniform go to the exact path)
    assert nok.value > len(S) // 4
