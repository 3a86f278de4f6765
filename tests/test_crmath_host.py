"""The device's correctly rounded log/log1p/sin/cos (mckl_amd/csrc/crmath.hpp),
compiled for the host from the same source, against quad-precision values
(libquadmath through the oracle) rounded once.

Pins: the double-double tables of log(j/1024) and sin/cos(j/64) to 2^-100
relative; log of every kind of U01OC double the Normal fill draws (2^20
random ones plus the extremes), of doubles across the whole exponent range
and next to 1; sin and cos of the Normal fill's angles 2*pi*u and of
arguments next to each multiple of pi/2 up to 2*pi; log1p across (-1, 1e300)
including both sides of its 2^-11 switch and the particle-filter example's
arguments 0.1 (10 d)^2.  Every result must equal
the correctly rounded value.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "crmath_host.cpp")
LIB = os.path.join(ROOT, "tests", "cpp", "libcrmath_host.so")
HDR = os.path.join(ROOT, "mckl_amd", "csrc", "crmath.hpp")


def build_host_lib():
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math",
           "-fPIC", "-shared", "-Wall", "-o", LIB, SRC]
    subprocess.run(cmd, check=True)


@pytest.fixture(scope="module")
def crm():
    if (not os.path.exists(LIB) or
            os.path.getmtime(LIB) < max(os.path.getmtime(SRC),
                                        os.path.getmtime(HDR))):
        build_host_lib()
    return ctypes.CDLL(LIB)


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _log(crm, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(len(x))
    crm.crm_host_log(ctypes.c_size_t(len(x)), P(x), P(y))
    return y


def _sincos(crm, x):
    x = np.ascontiguousarray(x, np.float64)
    s = np.zeros(len(x))
    c = np.zeros(len(x))
    crm.crm_host_sincos(ctypes.c_size_t(len(x)), P(x), P(s), P(c))
    return s, c


def test_tables(crm):
    nl, nt, j0 = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    crm.crm_host_sizes(ctypes.byref(nl), ctypes.byref(nt), ctypes.byref(j0))
    lc = np.zeros(2 * nl.value)
    sa = np.zeros(2 * nt.value)
    ca = np.zeros(2 * nt.value)
    crm.crm_host_tables(P(lc), P(sa), P(ca))
    c = np.arange(j0.value, j0.value + nl.value) / 1024.0
    nz = c != 1.0
    e = orc.quad_dd_relerr(orc.CR_LOG, c[nz], lc[0::2][nz], lc[1::2][nz])
    assert e.max() < 2.0 ** -100
    assert lc[2 * (1024 - j0.value)] == 0.0
    a = np.arange(nt.value) / 64.0
    e = orc.quad_dd_relerr(orc.CR_SIN, a[1:], sa[2::2], sa[3::2])
    assert e.max() < 2.0 ** -100
    e = orc.quad_dd_relerr(orc.CR_COS, a, ca[0::2], ca[1::2])
    assert e.max() < 2.0 ** -100


def test_log_u01oc_inputs(crm):
    rng = np.random.default_rng(2024)
    w = np.concatenate([rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64),
                        np.arange(0, 4096, dtype=np.uint64),
                        np.arange(2 ** 32 - 4096, 2 ** 32, dtype=np.uint64)])
    u = (w.astype(np.float64) + 1.0) * 2.0 ** -32  # U01OC: (w + 1) 2^-32
    got = _log(crm, u)
    want = orc.quad_eval(orc.CR_LOG, u)
    assert np.array_equal(got, want), int((got != want).sum())
    assert got[u == 1.0].tolist() == [0.0]


def test_log_general(crm):
    rng = np.random.default_rng(7)
    x = np.concatenate([
        np.exp(rng.uniform(-700.0, 700.0, 200000)),
        1.0 + rng.uniform(-2.0 ** -20, 2.0 ** -20, 50000),
        np.array([np.finfo(np.float64).tiny, np.finfo(np.float64).max, 0.75,
                  1.5, np.nextafter(1.5, 0.0), 2.0, 0.5])])
    got = _log(crm, x)
    want = orc.quad_eval(orc.CR_LOG, x)
    assert np.array_equal(got, want), int((got != want).sum())


def _log1p(crm, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(len(x))
    crm.crm_host_log1p(ctypes.c_size_t(len(x)), P(x), P(y))
    return y


def test_log1p(crm):
    rng = np.random.default_rng(11)
    d = rng.normal(0.0, 1.0, 200000) * np.exp(rng.uniform(-12, 3, 200000))
    t = 2.0 ** -11
    x = np.concatenate([
        0.1 * (10.0 * d) ** 2,                      # the PF example's form
        np.exp(rng.uniform(-700.0, 690.0, 100000)),
        -np.exp(rng.uniform(-700.0, -1e-12, 100000)),  # (-1, 0)
        t + np.arange(-300, 301) * 2.0 ** -63,         # both sides of 2^-11
        -t + np.arange(-300, 301) * 2.0 ** -63,
        np.nextafter(-1.0, 0.0) + np.arange(0, 64) * 2.0 ** -53,
        np.array([0.0, -0.0, 5e-324, -5e-324, 1.0, -0.5, 1e300])])
    got = _log1p(crm, x)
    want = orc.quad_eval(orc.CR_LOG1P, x)
    assert np.array_equal(got, want), int((got != want).sum())
    assert np.signbit(_log1p(crm, np.array([-0.0])))[0]


def test_sincos_normal_angles(crm):
    rng = np.random.default_rng(99)
    w = rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64)
    u = (w.astype(np.float64) + 1.0) * 2.0 ** -32
    th = 6.283185307179586 * u  # the fill's angles, rounded as it rounds
    s, c = _sincos(crm, th)
    assert np.array_equal(s, orc.quad_eval(orc.CR_SIN, th))
    assert np.array_equal(c, orc.quad_eval(orc.CR_COS, th))


def test_sincos_near_quadrant_boundaries(crm):
    # arguments next to k pi/2 (where sin or cos is tiny and the reduction
    # must keep relative accuracy) and next to the table's cells (j/64)
    xs = []
    for k in range(0, 5):
        base = k * np.pi / 2
        xs.append(base + np.arange(-500, 501) * 2.0 ** -50 * max(base, 1.0))
    for j in range(1, 52):
        for k in range(0, 4):
            xs.append(k * np.pi / 2 + j / 64.0 + np.arange(-20, 21) * 1e-17)
    x = np.concatenate(xs)
    x = x[(x > 0) & (x <= 8.0)]
    s, c = _sincos(crm, x)
    assert np.array_equal(s, orc.quad_eval(orc.CR_SIN, x))
    assert np.array_equal(c, orc.quad_eval(orc.CR_COS, x))


def test_check_detects_a_wrong_last_bit(crm):
    """The comparison itself can fail: one ulp off is caught."""
    x = np.array([0.3, 1.7, 2.9])
    want = orc.quad_eval(orc.CR_SIN, x)
    s, _ = _sincos(crm, x)
    assert np.array_equal(s, want)
    assert not np.array_equal(np.nextafter(s, 2.0), want)
