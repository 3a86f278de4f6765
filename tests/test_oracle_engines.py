"""Pin the oracle's engines: the reference's own KAT files, FIPS-197, and the
compiled reference (oracle/_ref) for ragged bulk calls, scalar calls and
discard.  CPU only."""
import ctypes
import os

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

ENGINES = [
    ("philox4x32", orc.PHILOX4X32, 32),
    ("philox4x32_64", orc.PHILOX4X32, 64),
    ("threefry4x64", orc.THREEFRY4X64, 32),
    ("threefry4x64_64", orc.THREEFRY4X64, 64),
    ("ars", orc.ARS, 32),
    ("ars_64", orc.ARS, 64),
]

needs_ref = pytest.mark.skipif(orc.REF is None,
                               reason="oracle/_ref not built here")


@pytest.mark.parametrize("name,kind,rbits", ENGINES)
def test_kat_files(name, kind, rbits):
    """example/random/include/random_rng.hpp:80-104: a default engine
    (key {1,0,..}, counter 0) bulk-fills exactly the KAT words."""
    kat = np.array([int(l) for l in open(os.path.join(
        GOLDEN, "random_%s.data" % name))], dtype=np.uint64)
    assert len(kat) == 1000
    e = orc.engine(kind, [1], rbits)
    got = orc.rand_u32(e, 1000) if rbits == 32 else orc.rand_u64(e, 1000)
    assert np.array_equal(got.astype(np.uint64), kat)


def test_aes_round_fips197():
    """FIPS-197 Appendix C.1 (the vector example/random/src/random_aes.cpp
    checks): AES-128 built from the same round function as ARS."""
    key = bytes(range(16))
    pt = bytes.fromhex("00112233445566778899aabbccddeeff")
    assert orc.aes128(key, pt).hex() == "69c4e0d86a7b0430d8cdb78070b4c55a"


def test_aes_tables_closed_form():
    t0 = (ctypes.c_uint32 * 256)()
    sb = (ctypes.c_uint8 * 256)()
    orc.O.mo_aes_tables(t0, sb)
    # S-box spot values (FIPS-197 Figure 7) and bijectivity
    assert sb[0x00] == 0x63 and sb[0x53] == 0xED and sb[0xFF] == 0x16
    assert len(set(sb)) == 256
    # T0 = (2S, S, S, 3S) little-endian; aes_constants.hpp:42 starts 0xA56363C6
    assert t0[0] == 0xA56363C6 and t0[1] == 0x847C7CF8


def test_philox_counter_words():
    """The 128-bit counter enters as four little-endian u32 words: a counter
    in the high word must differ from the same value in the low word."""
    a = orc.philox_block([1, 2], [5, 0])
    b = orc.philox_block([1, 2], [0, 5])
    assert a != b
    e = orc.engine(orc.PHILOX4X32, [1, 2])
    assert list(orc.rand_u32(e, 4)) == orc.philox_block([1, 2], [1, 0])


@needs_ref
@pytest.mark.parametrize("name,kind,rbits", ENGINES)
@pytest.mark.parametrize("seq", [
    [1, 1, 1, 7, 0, 0, 4096, 3],
    [5, -3, 0, 17, -1000, 2, 1],
    [0, 0, 0, 0, 0, 9, -7, 65],
    [12345, -2 ** 20, 77],
])
def test_engine_ops_vs_reference(name, kind, rbits, seq):
    key = [0xA5A5A5A5, 0x5A5A5A5A, 7, 9]
    ref, rctr, ridx = orc.ref_engine_ops(kind, rbits, key, seq)
    e = orc.engine(kind, key, rbits)
    out = []
    for s in seq:
        if s > 0:
            out.append(orc.rand_u32(e, s) if rbits == 32 else orc.rand_u64(e, s))
        elif s < 0:
            orc.O.mo_discard(ctypes.byref(e), ctypes.c_uint64(-s))
        else:
            out.append(orc.rand_u32(e, 1) if rbits == 32 else orc.rand_u64(e, 1))
    got = np.concatenate(out) if out else np.zeros(0)
    assert np.array_equal(got, ref)
    assert orc.state(e) == (rctr, ridx)


@needs_ref
@pytest.mark.parametrize("kind", [orc.PHILOX4X32, orc.THREEFRY4X64, orc.ARS])
def test_counter_carry_vs_reference(kind):
    """Bulk fills that cross 2^64 in the low counter word (the reference
    splits its SIMD loop there: random/internal/philox_sse2_4x32.hpp:71-92)."""
    key = [3, 1, 4, 1]
    start = [2 ** 64 - 5, 2 ** 64 - 1, 1, 0]
    ref, rctr, ridx = orc.ref_engine_ops(kind, 32, key, [3, 64, 1], ctr0=start)
    e = orc.engine(kind, key)
    c = (ctypes.c_uint64 * 4)(*start)
    orc.O.mo_engine_set_ctr(ctypes.byref(e), c)
    got = np.concatenate([orc.rand_u32(e, 3), orc.rand_u32(e, 64),
                          orc.rand_u32(e, 1)])
    assert np.array_equal(got, ref)
    assert orc.state(e) == (rctr, ridx)


@needs_ref
def test_discard_carry_quirk_vs_reference():
    """discard's increment(ctr, n) carries when ctr[0] == max - n exactly
    (increment_generic.hpp:208-217 tests with '<'): the oracle matches."""
    key = [1, 2]
    M = 4
    for lo in (2 ** 64 - 1 - 3, 2 ** 64 - 1 - 2, 2 ** 64 - 1):
        ref, rctr, ridx = orc.ref_engine_ops(orc.PHILOX4X32, 32, key,
                                             [-(3 * M + 1), 2], ctr0=[lo, 0])
        e = orc.engine(orc.PHILOX4X32, key)
        orc.O.mo_engine_set_ctr(ctypes.byref(e), (ctypes.c_uint64 * 4)(lo, 0))
        orc.O.mo_discard(ctypes.byref(e), ctypes.c_uint64(3 * M + 1))
        got = orc.rand_u32(e, 2)
        assert np.array_equal(got, ref) and orc.state(e) == (rctr, ridx)


def test_engine_text_fixture():
    """CounterEngine operator<< (random/counter.hpp:235-249) from the
    compiled reference after mixed call sequences (bulk, scalar, discard):
    the oracle's restatement prints the same text."""
    import json
    cases = json.load(open(os.path.join(GOLDEN, "engine_text.json")))
    for c in cases:
        e = orc.engine(c["kind"], c["key"], c["rbits"])
        for s in c["calls"]:
            if s > 0:
                if c["rbits"] == 32:
                    orc.rand_u32(e, s)
                else:
                    orc.rand_u64(e, s)
            elif s < 0:
                orc.O.mo_discard(ctypes.byref(e), ctypes.c_uint64(-s))
            else:
                if c["rbits"] == 32:
                    orc.O.mo_next_u32(ctypes.byref(e))
                else:
                    orc.rand_u64(e, 1)
        assert orc.engine_text(e) == c["text"], c


def _baseline_step(lib, n, nthreads):
    key = (ctypes.c_uint64 * 4)(0x0123456789ABCDEF, 0, 0, 0)
    u32 = np.zeros(n, np.uint32)
    f64 = np.zeros(n, np.float64)
    if nthreads is None:
        lib.ref_time_philox_u01(key, ctypes.c_size_t(n), 1, orc.P(u32),
                                orc.P(f64))
    else:
        lib.ref_time_philox_u01_mt(key, ctypes.c_size_t(n), 1, nthreads,
                                   orc.P(u32), orc.P(f64))
    return u32, f64


@needs_ref
@pytest.mark.parametrize("n,nthreads", [(1 << 16, 4), (1000, 3), (37, 8)])
def test_baseline_threads_match_single(n, nthreads):
    """The CPU baseline's threaded split of the full step (raw u32 fill plus
    U01OO doubles, each thread moved to its range by discard) writes the
    single-thread outputs, and the speed build (AVX-512 / AES-NI paths of the
    reference) writes the pinned build's."""
    u1, f1 = _baseline_step(orc.REF, n, None)
    libs = [orc.REF] + ([orc.REF_V4] if orc.REF_V4 is not None else [])
    for lib in libs:
        u, f = _baseline_step(lib, n, nthreads)
        assert np.array_equal(u, u1)
        assert np.array_equal(f.view(np.uint64), f1.view(np.uint64))
    e = orc.engine(orc.PHILOX4X32, [0x0123456789ABCDEF])
    assert np.array_equal(u1, orc.rand_u32(e, n))


@needs_ref
@pytest.mark.parametrize("kind", [orc.THREEFRY4X64, orc.ARS])
@pytest.mark.parametrize("dbl", [0, 1])
def test_baseline_normal_streams(kind, dbl):
    """The threaded Normal-streams baseline fills stream s with the
    reference's rand(NormalDistribution) from key {s + 1, 0, 0, 0}; the speed
    build agrees with the pinned one to a few ulp (its log/sin/cos and
    contraction may differ)."""
    ns, n = 5, 3001
    dt = np.float64 if dbl else np.float32
    mean, sd = (2.5, 0.3) if dbl else (0.0, 1.0)
    outs = []
    libs = [orc.REF] + ([orc.REF_V4] if orc.REF_V4 is not None else [])
    for lib in libs:
        out = np.zeros(ns * n, dt)
        t = lib.ref_time_normal_streams(kind, dbl, ns, n, mean, sd, 3,
                                        orc.P(out))
        assert t >= 0
        outs.append(out)
    for s in range(ns):
        want = orc.ref_normal(kind, [s + 1, 0, 0, 0], dt, n, mean, sd)[0]
        assert np.array_equal(outs[0][s * n:(s + 1) * n], want)
    if len(outs) > 1:
        a, b = outs
        assert np.allclose(a, b, rtol=8 * np.finfo(dt).eps, atol=0)
