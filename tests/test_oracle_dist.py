"""Pin the oracle's U01 and Normal fills: closed forms and edge words for U01,
the compiled reference (bit-exact) for both, the committed fixtures, and
moments against theory.  CPU only."""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(orc.REF is None,
                               reason="oracle/_ref not built here")


def test_u01_closed_forms_and_edges():
    """random/internal/u01_generic.hpp:48-164 at u = 0, 1, 2^31, 2^32-1."""
    O = orc.O
    for u in (0, 1, 255, 256, 511, 512, 2 ** 31, 2 ** 32 - 1):
        assert O.mo_u01_u32_f64(orc.CO, u) == u * 2.0 ** -32
        assert O.mo_u01_u32_f64(orc.OC, u) == u * 2.0 ** -32 + 2.0 ** -32
        assert O.mo_u01_u32_f64(orc.OO, u) == u * 2.0 ** -32 + 2.0 ** -33
        assert O.mo_u01_u32_f64(orc.CC, u) == (u + (u & 1)) * 2.0 ** -32
        f = np.float32
        assert O.mo_u01_u32_f32(orc.CO, u) == f((u >> 8) * 2.0 ** -24)
        assert O.mo_u01_u32_f32(orc.OC, u) == f(((u >> 8) + 1) * 2.0 ** -24)
        assert O.mo_u01_u32_f32(orc.OO, u) == f((u >> 9) * 2.0 ** -23 + 2.0 ** -24)
        v = ((u << 1) & 0xFFFFFFFF) >> 7
        assert O.mo_u01_u32_f32(orc.CC, u) == f((v + (v & 1)) * 2.0 ** -25)
    # interval bounds (example/random/include/random_u01.hpp:119-127)
    assert O.mo_u01_u32_f64(orc.OO, 0) > 0 and O.mo_u01_u32_f64(orc.OO, 2 ** 32 - 1) < 1
    assert O.mo_u01_u32_f64(orc.CC, 2 ** 32 - 1) == 1.0
    assert O.mo_u01_u32_f32(orc.OO, 2 ** 32 - 1) < 1.0
    assert O.mo_u01_u32_f32(orc.OC, 2 ** 32 - 1) == 1.0
    # Canonical<double>: u0 2^-64 + u1 2^-32, one rounding
    assert O.mo_u01_canonical_f64(0, 0) == 0.0
    assert O.mo_u01_canonical_f64(2 ** 32 - 1, 2 ** 32 - 1) == \
        float(np.float64((2 ** 64 - 1) * 2.0 ** -64))


def rn(u, bits):
    """u rounded to `bits` significant bits, to nearest, ties to even"""
    e = u.bit_length() - bits
    if e <= 0:
        return u
    q, r = divmod(u, 2 ** e)
    half = 2 ** (e - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return q * 2 ** e


def test_u01_u64_closed_forms_and_edges():
    """The W = 64 forms (u01_generic.hpp:48-198 with UIntType uint64, the
    word type of every real type on a 64-bit-result engine,
    u01_distribution.hpp:105-113) at edge words, against exact rational
    arithmetic."""
    from fractions import Fraction as F
    O = orc.O
    O.mo_u01_u64_f64.restype = ctypes.c_double
    O.mo_u01_u64_f32.restype = ctypes.c_float
    O.mo_u01_canonical64_f64.restype = ctypes.c_double
    O.mo_u01_canonical64_f32.restype = ctypes.c_float
    for f in (O.mo_u01_u64_f64, O.mo_u01_u64_f32):
        f.argtypes = [ctypes.c_int, ctypes.c_uint64]
    for f in (O.mo_u01_canonical64_f64, O.mo_u01_canonical64_f32):
        f.argtypes = [ctypes.c_uint64]
    words = (0, 1, 2 ** 11 - 1, 2 ** 11, 2 ** 40, 2 ** 41 - 1, 2 ** 53 - 1,
             2 ** 63, 2 ** 64 - 2 ** 11, 2 ** 64 - 1, 0x0123456789ABCDEF,
             2 ** 24 + 1, (2 ** 24 + 1) * 2 ** 39, (2 ** 24 + 3) * 2 ** 39,
             2 ** 53 + 1, (2 ** 53 + 1) * 2 ** 10, (2 ** 53 + 3) * 2 ** 10)
    for u in words:
        assert O.mo_u01_u64_f64(orc.CO, u) == F(u >> 11, 2 ** 53)
        assert O.mo_u01_u64_f64(orc.OC, u) == F((u >> 11) + 1, 2 ** 53)
        assert O.mo_u01_u64_f64(orc.OO, u) == F(2 * (u >> 12) + 1, 2 ** 53)
        v = ((u << 1) % 2 ** 64) >> 10
        assert O.mo_u01_u64_f64(orc.CC, u) == F(v + (v & 1), 2 ** 54)
        assert O.mo_u01_u64_f32(orc.CO, u) == F(u >> 40, 2 ** 24)
        assert O.mo_u01_u64_f32(orc.OC, u) == F((u >> 40) + 1, 2 ** 24)
        assert O.mo_u01_u64_f32(orc.OO, u) == F(2 * (u >> 41) + 1, 2 ** 24)
        v = ((u << 1) % 2 ** 64) >> 39
        assert O.mo_u01_u64_f32(orc.CC, u) == F(v + (v & 1), 2 ** 25)
        # Canonical: the integer rounded once to the real type, times 2^-64
        assert O.mo_u01_canonical64_f64(u) == F(rn(u, 53), 2 ** 64)
        assert O.mo_u01_canonical64_f32(u) == F(rn(u, 24), 2 ** 64)
    assert O.mo_u01_u64_f64(orc.OO, 2 ** 64 - 1) < 1.0
    assert O.mo_u01_u64_f64(orc.OC, 2 ** 64 - 1) == 1.0
    assert O.mo_u01_u64_f64(orc.CC, 2 ** 64 - 1) == 1.0
    assert O.mo_u01_u64_f32(orc.OO, 2 ** 64 - 1) < 1.0


def _pre(e, pre, rbits):
    """pre raw result words (u32 or u64) before the distribution call"""
    if rbits == 64:
        orc.rand_u64(e, pre)
    else:
        orc.rand_u32(e, pre)


@needs_ref
@pytest.mark.parametrize("rbits", [32, 64])
@pytest.mark.parametrize("kind", [orc.PHILOX4X32, orc.THREEFRY4X64, orc.ARS])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("interval", [orc.CC, orc.CO, orc.OC, orc.OO,
                                      orc.CANONICAL])
def test_u01_vs_reference(kind, dtype, interval, rbits):
    for pre, n in ((0, 5000), (3, 2049), (1, 1), (7, 0), (2, 3000)):
        ref, rctr, ridx = orc.ref_u01(kind, [5], interval, dtype, n, pre,
                                      rbits)
        e = orc.engine(kind, [5], rbits)
        _pre(e, pre, rbits)
        got = orc.rand_u01(e, interval, n, dtype)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
        assert orc.state(e) == (rctr, ridx)


UNIFORM_CASES = ((0, 5000, 0.0, 1.0), (3, 2049, -2.5, 7.25),
                 (1, 1, 1e-3, 1e-3 + 1e-9), (0, 4097, -1e6, -3.0),
                 (7, 0, 0.0, 1.0), (2, 3001, 100.0, 1e8))


@needs_ref
@pytest.mark.parametrize("rbits", [32, 64])
@pytest.mark.parametrize("kind", [orc.PHILOX4X32, orc.THREEFRY4X64, orc.ARS])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uniform_real_vs_reference(kind, dtype, rbits):
    """UniformRealDistribution (uniform_real_distribution.hpp:48-95) bit-exact
    against the compiled reference: chunked U01CO, r * (b - a) + a rounded
    twice, engine state after the call."""
    for pre, n, a, b in UNIFORM_CASES:
        ref, rctr, ridx = orc.ref_uniform_real(kind, [6], dtype, n, a, b, pre,
                                               rbits)
        e = orc.engine(kind, [6], rbits)
        _pre(e, pre, rbits)
        got = orc.rand_uniform_real(e, n, a, b, dtype)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), \
            (pre, n, a, b)
        assert orc.state(e) == (rctr, ridx)
        if n:
            assert got.min() >= dtype(a) and got.max() <= dtype(b)


@needs_ref
@pytest.mark.parametrize("rbits", [32, 64])
@pytest.mark.parametrize("kind", [orc.PHILOX4X32, orc.THREEFRY4X64, orc.ARS])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_normal_vs_reference(kind, dtype, rbits):
    """Chunked Box-Muller (normal_distribution.hpp:50-101) bit-exact, including
    partial chunks, the odd tail and an unaligned start."""
    for pre, n, mean, sd in ((0, 8192, 0.0, 1.0), (1, 5001, 2.5, 0.3),
                             (0, 1, 0.0, 1.0), (3, 4097, 1.0, 2.0),
                             (0, 2048 * 3 + 7, 0.0, 1.0), (2, 1025, -1.0, 0.5)):
        ref, rctr, ridx = orc.ref_normal(kind, [9], dtype, n, mean, sd, pre,
                                         rbits)
        e = orc.engine(kind, [9], rbits)
        _pre(e, pre, rbits)
        got = orc.rand_normal(e, n, mean, sd, dtype)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8)), (pre, n)
        assert orc.state(e) == (rctr, ridx)


def test_config1_fixture():
    """The committed config-1 hashes (written from the compiled reference)
    are reproduced by the oracle."""
    cfg = json.load(open(os.path.join(GOLDEN, "config1.json")))
    e = orc.engine(orc.PHILOX4X32, cfg["key"])
    a = orc.rand_u32(e, cfg["n"])
    assert hashlib.sha256(a.tobytes()).hexdigest() == cfg["sha256_u32"]
    assert orc.state(e) == (cfg["state_after_u32"]["ctr"],
                            cfg["state_after_u32"]["index"])
    e = orc.engine(orc.PHILOX4X32, cfg["key"])
    f = orc.rand_u01(e, orc.OO, cfg["n"], np.float64)
    assert hashlib.sha256(f.tobytes()).hexdigest() == cfg["sha256_u01oo_f64"]


@pytest.mark.parametrize("ename,kind", [("threefry", orc.THREEFRY4X64),
                                        ("ars", orc.ARS)])
def test_normal_stream_fixture(ename, kind):
    """Windows of the config-2 streams (seed 2, 2^22 values) from the
    compiled reference equal the oracle's."""
    fx = np.load(os.path.join(GOLDEN, "normal_streams.npz"))
    n = 1 << 22
    for dname, dt, mean, sd in (("f32", np.float32, 0.0, 1.0),
                                ("f64", np.float64, 2.5, 0.3)):
        e = orc.engine(kind, [2])
        r = orc.rand_normal(e, n, mean, sd, dt)
        p = "%s_%s_2" % (ename, dname)
        assert np.array_equal(r[:4096], fx[p + "_head"])
        assert np.array_equal(r[n // 2 - 2048:n // 2 + 2048], fx[p + "_mid"])
        assert np.array_equal(r[-4096:], fx[p + "_tail"])
        st = fx[p + "_state"]
        assert orc.state(e) == ([int(v) for v in st[:-1]], int(st[-1]))


def test_normal_moments():
    e = orc.engine(orc.THREEFRY4X64, [1])
    n = 1 << 20
    z = orc.rand_normal(e, n, 0.0, 1.0, np.float64)
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.var() - 1) < 5 * np.sqrt(2.0 / n)
    # fourth moment of N(0,1) is 3
    assert abs((z ** 4).mean() - 3) < 0.05


@needs_ref
@pytest.mark.parametrize("rbits", [32, 64])
@pytest.mark.parametrize("kind", [orc.PHILOX4X32, orc.THREEFRY4X64, orc.ARS])
@pytest.mark.parametrize("ubits", [16, 32, 64])
def test_uniform_bits_vs_reference(kind, rbits, ubits):
    """rand(rng, UniformBitsDistribution<U>, n, r) for U of 16, 32, 64 bits
    (uniform_bits_distribution.hpp:60-107) bit-exact against the compiled
    reference, engine state included, across chunk boundaries (4096 u16 or
    2048 u32 per chunk when narrower than the result word)."""
    for pre, n in ((0, 5000), (3, 4097), (1, 1), (7, 0), (2, 9001)):
        ref, rctr, ridx = orc.ref_uniform_bits(kind, [8], ubits, n, pre,
                                               rbits)
        e = orc.engine(kind, [8], rbits)
        _pre(e, pre, rbits)
        got = orc.rand_uniform_bits(e, ubits, n)
        assert np.array_equal(got, ref), (pre, n)
        assert orc.state(e) == (rctr, ridx)


@pytest.mark.parametrize("rbits", [32, 64])
def test_uniform_bits_closed_form(rbits):
    """The words themselves: u64 from a 32-bit engine is two consecutive
    result words, low first; narrower values are the low bits of one word
    each (the reference's scalar UniformBits, :128-175, gives the same)."""
    w = orc.rand_u32(orc.engine(orc.PHILOX4X32, [4], 32), 40) \
        if rbits == 32 else \
        orc.rand_u64(orc.engine(orc.PHILOX4X32, [4], 64), 20)
    e = orc.engine(orc.PHILOX4X32, [4], rbits)
    u16 = orc.rand_uniform_bits(e, 16, 20)
    assert np.array_equal(u16, (w[:20] & 0xFFFF).astype(np.uint16))
    if rbits == 32:
        e = orc.engine(orc.PHILOX4X32, [4], 32)
        u64 = orc.rand_uniform_bits(e, 64, 20)
        want = w[0::2].astype(np.uint64) | (w[1::2].astype(np.uint64) << 32)
        assert np.array_equal(u64, want)
    else:
        e = orc.engine(orc.PHILOX4X32, [4], 64)
        u32 = orc.rand_uniform_bits(e, 32, 20)
        assert np.array_equal(u32, (w & 0xFFFFFFFF).astype(np.uint32))
