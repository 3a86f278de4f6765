"""mckl_amd -- MI355X (gfx950) implementation of MCKL's Monte Carlo hot path.

Python mirror of the reference's C++ interface for this path, over the C ABI
in include/mckl_hip.h (libmckl_hip.so).  Names follow the reference:
``Philox4x32``, ``Threefry4x64``, ``ARS`` (and ``*_64``), ``rand(rng, n, r)``,
``rand(rng, dist, n, r)``, ``U01OODistribution``, ``NormalDistribution``,
``Weight``, ``ResampleSystematic`` ...  Output arrays are torch tensors in
device memory; torch supplies allocation and the current HIP stream only.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib, MCKLError  # noqa: F401

__all__ = [
    "Philox4x32", "Philox4x32_64", "Threefry4x64", "Threefry4x64_64", "ARS",
    "ARS_64", "rand", "U01CCDistribution", "U01CODistribution",
    "U01OCDistribution", "U01OODistribution", "U01CanonicalDistribution",
    "NormalDistribution", "MCKLError",
]


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class _CounterEngine:
    """CounterEngine<ResultType, Generator> (random/counter.hpp:61-288)."""

    KIND = None
    RBITS = 32
    KEY_WORDS = 2

    def __init__(self, s=1, key=None):
        self.state = _lib.Engine()
        if key is not None:
            self.seed_key(key)
        else:
            k = (ctypes.c_uint64 * 4)(0, 0, 0, 0)
            check(lib.mckl_hip_engine_init(ctypes.byref(self.state), self.KIND,
                                           self.RBITS, k))
            self.seed(s)

    # seed(result_type s): counter.hpp:97-103
    def seed(self, s):
        check(lib.mckl_hip_engine_seed(ctypes.byref(self.state), int(s)))

    # seed(const key_type&): counter.hpp:112-114
    def seed_key(self, key):
        k = list(key) + [0] * (4 - len(key))
        arr = (ctypes.c_uint64 * 4)(*[int(v) for v in k])
        check(lib.mckl_hip_engine_init(ctypes.byref(self.state), self.KIND,
                                       self.RBITS, arr))

    def key(self):
        return [int(v) for v in self.state.key[:self.KEY_WORDS]]

    def ctr(self):
        n = 4 if self.KIND == _lib.THREEFRY4X64 else 2
        return [int(v) for v in self.state.ctr[:n]]

    def set_ctr(self, c):
        arr = (ctypes.c_uint64 * 4)(*(list(c) + [0] * (4 - len(c))))
        check(lib.mckl_hip_engine_set_ctr(ctypes.byref(self.state), arr))

    @property
    def index(self):
        return int(self.state.index)

    def discard(self, nskip):
        check(lib.mckl_hip_engine_discard(ctypes.byref(self.state), int(nskip)))

    def __call__(self):
        out = ctypes.c_uint64()
        check(lib.mckl_hip_engine_next(ctypes.byref(self.state),
                                       ctypes.byref(out)))
        return int(out.value)

    def copy(self):
        other = object.__new__(type(self))
        other.state = _lib.Engine()
        ctypes.pointer(other.state)[0] = self.state
        return other

    def __eq__(self, other):
        return (type(self) is type(other)
                and bytes(self.state) == bytes(other.state))

    @property
    def words(self):
        return lib.mckl_hip_engine_words(ctypes.byref(self.state))


class Philox4x32(_CounterEngine):
    KIND, RBITS, KEY_WORDS = _lib.PHILOX4X32, 32, 2


class Philox4x32_64(_CounterEngine):
    KIND, RBITS, KEY_WORDS = _lib.PHILOX4X32, 64, 2


class Threefry4x64(_CounterEngine):
    KIND, RBITS, KEY_WORDS = _lib.THREEFRY4X64, 32, 4


class Threefry4x64_64(_CounterEngine):
    KIND, RBITS, KEY_WORDS = _lib.THREEFRY4X64, 64, 4


class ARS(_CounterEngine):
    KIND, RBITS, KEY_WORDS = _lib.ARS, 32, 4


class ARS_64(_CounterEngine):
    KIND, RBITS, KEY_WORDS = _lib.ARS, 64, 4


class _U01:
    INTERVAL = None

    def __init__(self, dtype="float64"):
        self.dtype = np.dtype(dtype)

    def fill(self, rng, n, r):
        fn = (lib.mckl_hip_rand_u01_f64 if self.dtype == np.float64
              else lib.mckl_hip_rand_u01_f32)
        check(fn(ctypes.byref(rng.state), self.INTERVAL, n, _ptr(r), _stream()))


class U01CCDistribution(_U01):
    INTERVAL = _lib.U01_CC


class U01CODistribution(_U01):
    INTERVAL = _lib.U01_CO


class U01OCDistribution(_U01):
    INTERVAL = _lib.U01_OC


class U01OODistribution(_U01):
    INTERVAL = _lib.U01_OO


class U01CanonicalDistribution(_U01):
    INTERVAL = _lib.U01_CANONICAL


class NormalDistribution:
    """NormalDistribution<RealType> (random/normal_distribution.hpp:113-154)."""

    def __init__(self, mean=0.0, stddev=1.0, dtype="float64"):
        self.mean = mean
        self.stddev = stddev
        self.dtype = np.dtype(dtype)

    def fill(self, rng, n, r):
        if self.dtype == np.float64:
            check(lib.mckl_hip_rand_normal_f64(
                ctypes.byref(rng.state), n, self.mean, self.stddev, _ptr(r),
                _stream()))
        else:
            check(lib.mckl_hip_rand_normal_f32(
                ctypes.byref(rng.state), n, self.mean, self.stddev, _ptr(r),
                _stream()))


def rand(rng, *args):
    """rand(rng, n, r) or rand(rng, dist, n, r); r is a device tensor.

    random/counter.hpp:298-303 and random/internal/common.hpp:706-712.
    """
    if len(args) == 2:
        n, r = args
        if r.numel() < n:
            raise ValueError("output tensor too small")
        check(lib.mckl_hip_rand(ctypes.byref(rng.state), n, _ptr(r), _stream()))
        return r
    dist, n, r = args
    if r.numel() < n:
        raise ValueError("output tensor too small")
    dist.fill(rng, n, r)
    return r


def streams_normal(engines, n, mean, stddev, r):
    """One launch filling len(engines) streams of n Normal values each."""
    arr = (_lib.Engine * len(engines))()
    for i, e in enumerate(engines):
        arr[i] = e.state
    if r.dtype.itemsize == 8:
        fn = lib.mckl_hip_rand_normal_f64_streams
    else:
        fn = lib.mckl_hip_rand_normal_f32_streams
    check(fn(arr, len(engines), n, mean, stddev, _ptr(r), _stream()))
    for i, e in enumerate(engines):
        e.state = arr[i]
    return r
