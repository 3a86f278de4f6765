"""mckl_amd -- MI355X (gfx950) implementation of MCKL's Monte Carlo hot path.

Python mirror of the reference's C++ interface for this path, over the C ABI
in include/mckl_hip.h (libmckl_hip.so).  Names follow the reference:
``Philox4x32``, ``Threefry4x64``, ``ARS`` (and ``*_64``), ``rand(rng, n, r)``,
``rand(rng, dist, n, r)``, ``U01OODistribution``, ``NormalDistribution``,
``Weight``, ``ResampleSystematic`` ...  Output arrays are torch tensors in
device memory; torch supplies allocation and the current HIP stream only.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check, lib, MCKLError  # noqa: F401

__all__ = [
    "Philox4x32", "Philox4x32_64", "Threefry4x64", "Threefry4x64_64", "ARS",
    "ARS_64", "rand", "U01CCDistribution", "U01CODistribution",
    "U01OCDistribution", "U01OODistribution", "U01CanonicalDistribution",
    "NormalDistribution", "UniformRealDistribution",
    "UniformBitsDistribution", "MCKLError",
    "Particle", "ResampleEval", "SMCEstimator", "GenericSMCSampler",
]


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _init_device():
    """Initialise torch's HIP context before the library's first device call
    (a handle allocates device memory when it is created); torch cannot
    attach to the device afterwards on this runtime."""
    import torch
    torch.cuda.init()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


_FLOAT = {8: "float64", 4: "float32"}


def _dev(t, n, itemsize, what, floating=None):
    """Device pointer of tensor t after checking it: a contiguous tensor in
    GPU memory with elements of `itemsize` bytes (floating point when
    `floating` is True, integer when False) holding at least n of them.
    Raises MCKLError before any launch otherwise."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise MCKLError("%s: expected a torch tensor" % what)
    if not t.is_cuda:
        raise MCKLError("%s: tensor is not in GPU memory" % what)
    if not t.is_contiguous():
        raise MCKLError("%s: tensor is not contiguous" % what)
    if itemsize is not None and t.element_size() != itemsize:
        raise MCKLError("%s: %d-byte elements, expected %d" %
                        (what, t.element_size(), itemsize))
    if floating is not None and t.dtype.is_floating_point != floating:
        raise MCKLError("%s: expected %s elements, got %s" %
                        (what, _FLOAT.get(itemsize, "floating") if floating
                         else "integer", t.dtype))
    if t.numel() < n:
        raise MCKLError("%s: holds %d values, needs %d" % (what, t.numel(), n))
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

    # operator<< (random/counter.hpp:235-249): result_, ctr_, generator_,
    # index_, each array as "size v0 v1 ..."; the buffered block result_ is
    # recomputed on the device from (key, ctr)
    def state_string(self):
        st = self.state
        nbytes = 32 if self.KIND == _lib.THREEFRY4X64 else 16
        blk = (ctypes.c_uint8 * nbytes)()
        check(lib.mckl_hip_engine_block(ctypes.byref(st), blk))
        raw = bytes(blk)
        rb = self.RBITS // 8
        res = [int.from_bytes(raw[i:i + rb], "little")
               for i in range(0, nbytes, rb)]
        K = 4 if self.KIND == _lib.THREEFRY4X64 else 2
        parts = [str(len(res))] + [str(v) for v in res]
        parts += [str(K)] + [str(int(v)) for v in st.ctr[:K]]
        key = [int(v) for v in st.key]
        if self.KIND == _lib.PHILOX4X32:
            parts += ["2"] + [str(v & 0xFFFFFFFF) for v in key[:2]]
        elif self.KIND == _lib.ARS:
            parts += ["4"] + [str(v & 0xFFFFFFFF) for v in key]
        else:
            par = 0x1BD11BDAA9FC1A22
            for v in key:
                par ^= v
            parts += ["8"] + [str(v) for v in key] + [str(par), "0", "0", "0"]
        parts.append(str(int(st.index)))
        return " ".join(parts)

    # operator>> (random/counter.hpp:251-271)
    def load_state_string(self, text):
        v = [int(x) for x in text.split()]
        m = v[0]
        pos = 1 + m
        K = v[pos]
        ctr = v[pos + 1:pos + 1 + K]
        pos += 1 + K
        nk = v[pos]
        key = v[pos + 1:pos + 1 + nk][:self.KEY_WORDS]
        pos += 1 + nk
        index = v[pos]
        if len(v) != pos + 1 or index > m:
            raise MCKLError("malformed engine state")
        self.seed_key(key)
        self.set_ctr(ctr)
        self.state.index = index


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
        p = _dev(r, n, self.dtype.itemsize, "U01 output", True)
        check(fn(ctypes.byref(rng.state), self.INTERVAL, n, p, _stream()))


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
        p = _dev(r, n, self.dtype.itemsize, "Normal output", True)
        if self.dtype == np.float64:
            check(lib.mckl_hip_rand_normal_f64(
                ctypes.byref(rng.state), n, self.mean, self.stddev, p,
                _stream()))
        else:
            check(lib.mckl_hip_rand_normal_f32(
                ctypes.byref(rng.state), n, self.mean, self.stddev, p,
                _stream()))


class UniformBitsDistribution:
    """UniformBitsDistribution<UIntType> for UIntType of 16, 32 or 64 bits
    (random/uniform_bits_distribution.hpp:120-210): the engine's result
    words themselves when UIntType is at least as wide, else the low bits of
    one result word per value."""

    def __init__(self, bits=32):
        if bits not in (16, 32, 64):
            raise MCKLError("UniformBitsDistribution needs 16, 32 or 64 bits")
        self.bits = int(bits)

    def fill(self, rng, n, r):
        p = _dev(r, n, self.bits // 8, "UniformBits output", False)
        check(lib.mckl_hip_rand_uniform_bits(ctypes.byref(rng.state),
                                             self.bits, n, p, _stream()))


class UniformRealDistribution:
    """UniformRealDistribution<RealType> (random/uniform_real_distribution.hpp:
    65-95): U01CO scaled to [a, b) as r * (b - a) + a, two roundings."""

    def __init__(self, a=0.0, b=1.0, dtype="float64"):
        if not a < b:
            raise MCKLError("**UniformRealDistribution** constructed with "
                            "invalid arguments")
        self.a = a
        self.b = b
        self.dtype = np.dtype(dtype)

    def min(self):
        return self.a

    def max(self):
        return self.b

    def fill(self, rng, n, r):
        fn = (lib.mckl_hip_rand_uniform_real_f64 if self.dtype == np.float64
              else lib.mckl_hip_rand_uniform_real_f32)
        p = _dev(r, n, self.dtype.itemsize, "UniformReal output", True)
        check(fn(ctypes.byref(rng.state), n, self.a, self.b, p, _stream()))


def rand(rng, *args):
    """rand(rng, n, r) or rand(rng, dist, n, r); r is a device tensor.

    random/counter.hpp:298-303 and random/internal/common.hpp:706-712.
    """
    if len(args) == 2:
        n, r = args
        p = _dev(r, n, rng.RBITS // 8, "rand output", False)
        check(lib.mckl_hip_rand(ctypes.byref(rng.state), n, p, _stream()))
        return r
    dist, n, r = args
    dist.fill(rng, n, r)
    return r


def streams_normal(engines, n, mean, stddev, r):
    """One launch filling len(engines) streams of n Normal values each."""
    arr = (_lib.Engine * len(engines))()
    for i, e in enumerate(engines):
        arr[i] = e.state
    p = _dev(r, len(engines) * n, r.element_size(), "Normal output", True)
    if r.element_size() == 8:
        fn = lib.mckl_hip_rand_normal_f64_streams
    else:
        fn = lib.mckl_hip_rand_normal_f32_streams
    check(fn(arr, len(engines), n, mean, stddev, p, _stream()))
    for i, e in enumerate(engines):
        e.state = arr[i]
    return r


# ---------------------------------------------------------------------------
# Weight and resampling (core/weight.hpp, algorithm/resample.hpp)
# ---------------------------------------------------------------------------
MULTINOMIAL = _lib.MULTINOMIAL
STRATIFIED = _lib.STRATIFIED
SYSTEMATIC = _lib.SYSTEMATIC
RESIDUAL = _lib.RESIDUAL
RESIDUAL_STRATIFIED = _lib.RESIDUAL_STRATIFIED
RESIDUAL_SYSTEMATIC = _lib.RESIDUAL_SYSTEMATIC
MODEL_LG = _lib.MODEL_LG
MODEL_SV = _lib.MODEL_SV


class Weight:
    """Device-resident normalised FP64 weights (core/weight.hpp:44-228).

    ``log_nc()`` returns the log normaliser increment lmax + log(accw) of the
    last set_log/add_log, which the reference computes and discards.
    """

    def __init__(self, n):
        import torch
        self.data = torch.empty(n, dtype=torch.float64, device="cuda")
        self._stats = torch.zeros(2, dtype=torch.float64, device="cuda")
        self.set_equal()

    def size(self):
        return self.data.numel()

    def set_equal(self):
        check(lib.mckl_hip_weight_set_equal(self.size(), _ptr(self.data),
                                            _ptr(self._stats), _stream()))

    def set_log(self, v):
        p = _dev(v, self.size(), 8, "Weight.set_log values", True)
        check(lib.mckl_hip_weight_set_log(self.size(), p, _ptr(self.data),
                                          _ptr(self._stats), _stream()))

    def add_log(self, v):
        p = _dev(v, self.size(), 8, "Weight.add_log values", True)
        check(lib.mckl_hip_weight_add_log(self.size(), p, _ptr(self.data),
                                          _ptr(self._stats), _stream()))

    def stats(self):
        """The device tensor {ESS, log normaliser increment} of the last
        weight operation (read it without synchronising)."""
        return self._stats

    def ess(self):
        return float(self._stats[0].item())

    def log_nc(self):
        return float(self._stats[1].item())


def u01_rand_sequence(scheme, rng, n, r):
    p = _dev(r, n, 8, "uniform sequence output", True)
    check(lib.mckl_hip_u01_rand_sequence(scheme, ctypes.byref(rng.state), n,
                                         p, _stream()))
    return r


def resample_trans_u01_rep(N, M, w, u, rep):
    pw = _dev(w, N, 8, "weights", True)
    pu = _dev(u, M, 8, "uniforms", True)
    pr = _dev(rep, N, 4, "replication numbers", False)
    check(lib.mckl_hip_resample_trans_u01_rep(N, M, pw, pu, pr, _stream()))
    return rep


def resample_trans_rep_index(N, M, rep, idx):
    pr = _dev(rep, N, 4, "replication numbers", False)
    pi = _dev(idx, M, 4, "parent indices", False)
    check(lib.mckl_hip_resample_trans_rep_index(N, M, pr, pi, _stream()))
    return idx


class _ResampleAlgorithm:
    """ResampleAlgorithm<U01Seq, false> (algorithm/resample.hpp:515-576):
    ``alg(N, M, rng, w, rep, idx=None)`` on device arrays."""

    SCHEME = None

    def __call__(self, N, M, rng, w, rep, idx=None):
        pw = _dev(w, N, 8, "weights", True)
        pr = _dev(rep, N, 4, "replication numbers", False)
        pi = (_dev(idx, M, 4, "parent indices", False) if idx is not None
              else None)
        check(lib.mckl_hip_resample(self.SCHEME, N, M, ctypes.byref(rng.state),
                                    pw, pr, pi, _stream()))
        return rep


class ResampleMultinomial(_ResampleAlgorithm):
    SCHEME = MULTINOMIAL


class ResampleStratified(_ResampleAlgorithm):
    SCHEME = STRATIFIED


class ResampleSystematic(_ResampleAlgorithm):
    SCHEME = SYSTEMATIC


# ResampleAlgorithm<U01Seq, true>: residual family (resample.hpp:581-593)
class ResampleResidual(_ResampleAlgorithm):
    SCHEME = RESIDUAL


class ResampleResidualStratified(_ResampleAlgorithm):
    SCHEME = RESIDUAL_STRATIFIED


class ResampleResidualSystematic(_ResampleAlgorithm):
    SCHEME = RESIDUAL_SYSTEMATIC


def estimate(u, w, dim, layout=None, out=None):
    """SMCEstimator::estimate's weighted sums (algorithm/smc.hpp:89-118):
    out[j] = sum_i w[i] u(i, j) for u holding n x dim values (float64 or
    float32) in RowMajor (u[i*dim + j], the default) or ColMajor layout;
    w float64.  Returns a float64 device tensor of dim values."""
    import torch
    layout = _lib.ROW_MAJOR if layout is None else layout
    n = w.numel()
    pw = _dev(w, n, 8, "estimate weights", True)
    pu = _dev(u, n * dim, u.element_size() if u.element_size() in (4, 8)
              else 8, "estimate values", True)
    if out is None:
        out = torch.empty(max(dim, 1), dtype=torch.float64, device=w.device)
    po = _dev(out, dim, 8, "estimate output", True)
    fn = (lib.mckl_hip_estimate_f64 if u.dtype == torch.float64
          else lib.mckl_hip_estimate_f32)
    check(fn(n, dim, layout, pu, pw, po, _stream()))
    return out[:dim]


def select(idx, src, dst, dim=1):
    """StateMatrix<RowMajor, dim, float>::select, out of place."""
    M = idx.numel()
    pi = _dev(idx, M, 4, "select indices", False)
    ps = _dev(src, 0, 4, "select source", True)
    pd = _dev(dst, M * dim, 4, "select destination", True)
    check(lib.mckl_hip_select_f32(M, dim, pi, ps, pd, _stream()))
    return dst


RowMajor = _lib.ROW_MAJOR
ColMajor = _lib.COL_MAJOR
_SIZE_MAX = 2 ** 64 - 1


class ResampleIndex:
    """ResampleIndex<uint32_t> (algorithm/resample.hpp:597-832) with the
    history in device memory.  Index vectors are int32/uint32 device tensors
    (u32 entries); the trace-back runs on the device.  Iteration defaults
    follow the reference: ``iter_back`` = num_iter() - 1, ``iter`` = 0."""

    def __init__(self):
        h = ctypes.c_void_p()
        _init_device()
        check(lib.mckl_hip_resample_index_create(ctypes.byref(h)))
        self._h = h

    # the library's destroy function, bound while the module is alive (at
    # interpreter exit the module's globals may already be cleared)
    _destroy = staticmethod(lib.mckl_hip_resample_index_destroy)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._destroy(h)
            self._h = None

    def num_iter(self):
        n = ctypes.c_size_t()
        check(lib.mckl_hip_resample_index_num_iter(self._h, ctypes.byref(n)))
        return n.value

    def size(self, iter=None):
        n = ctypes.c_size_t()
        check(lib.mckl_hip_resample_index_size(
            self._h, _SIZE_MAX if iter is None else iter, ctypes.byref(n)))
        return n.value

    def reset(self):
        check(lib.mckl_hip_resample_index_reset(self._h))

    def clear(self):
        check(lib.mckl_hip_resample_index_clear(self._h))

    def push_back(self, N, idx=None):
        """push_back(N) (identity) or push_back(N, idx) with idx a device
        tensor of at least N 32-bit entries."""
        p = (_dev(idx, N, 4, "push_back indices", False) if idx is not None
             else None)
        check(lib.mckl_hip_resample_index_push_back(self._h, N, p, _stream()))

    def _back(self, iter_back):
        return (self.num_iter() - 1) % (2 ** 64) if iter_back is None \
            else iter_back

    def index(self, id, iter_back=None, iter=0):
        """index(id, iter_back, iter): an int for an int id; for a device
        tensor of ids, a device tensor of traced indices."""
        import torch
        ib = self._back(iter_back)
        scalar = not torch.is_tensor(id)
        ids = torch.tensor([id], dtype=torch.int32, device="cuda") \
            if scalar else id
        pid = _dev(ids, ids.numel(), 4, "index ids", False)
        out = torch.empty(ids.numel(), dtype=torch.int32, device=ids.device)
        check(lib.mckl_hip_resample_index_index(
            self._h, ids.numel(), pid, ib, iter, _ptr(out), _stream()))
        if scalar:
            return int(out.cpu().numpy().view(np.uint32)[0])
        return out

    def _shape(self, iter_back, iter):
        nr, nc = ctypes.c_size_t(), ctypes.c_size_t()
        check(lib.mckl_hip_resample_index_matrix_shape(
            self._h, iter_back, iter, ctypes.byref(nr), ctypes.byref(nc)))
        return nr.value, nc.value

    def index_matrix_nrow(self, iter_back=None):
        return self._shape(self._back(iter_back), 0)[0]

    def index_matrix_ncol(self, iter_back=None, iter=0):
        return self._shape(self._back(iter_back), iter)[1]

    def read_index_matrix(self, layout, out, iter_back=None, iter=0):
        ib = self._back(iter_back)
        nr, nc = self._shape(ib, iter)
        p = _dev(out, nr * nc, 4, "read_index_matrix output", False)
        check(lib.mckl_hip_resample_index_read_matrix(
            self._h, layout, ib, iter, p, _stream()))
        return out

    def index_matrix(self, layout, iter_back=None, iter=0):
        import torch
        ib = self._back(iter_back)
        nr, nc = self._shape(ib, iter)
        out = torch.empty(max(nr * nc, 1), dtype=torch.int32, device="cuda")
        self.read_index_matrix(layout, out, ib, iter)
        return out[:nr * nc]


# ---------------------------------------------------------------------------
# SMC samplers (algorithm/smc.hpp:151-421 shape) for the built-in models
# ---------------------------------------------------------------------------
def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class SMCSampler:
    """Bootstrap particle filter on the device.

    model: MODEL_LG (theta = a, sx, sy) or MODEL_SV (theta = phi, sigma,
    beta); y: host observations (length T).  ``iterate()`` enqueues all T
    steps on the current stream without host synchronisation.
    """

    def __init__(self, model, theta, y, N, seed=1, resample_threshold=0.5,
                 scheme=SYSTEMATIC):
        self._th = np.ascontiguousarray(theta, np.float64)
        self._y = np.ascontiguousarray(y, np.float64)
        self.T = len(self._y)
        self.N = int(N)
        h = ctypes.c_void_p()
        _init_device()
        check(lib.mckl_hip_smc_create(model, _dptr(self._th), self.T,
                                      _dptr(self._y), self.N, int(seed),
                                      float(resample_threshold), scheme,
                                      ctypes.byref(h)))
        self._h = h

    def iterate(self):
        check(lib.mckl_hip_smc_run(self._h, _stream()))

    def results(self):
        ll = ctypes.c_double()
        ess = np.zeros(self.T, np.float64)
        res = np.zeros(self.T, np.int32)
        check(lib.mckl_hip_smc_result(
            self._h, ctypes.byref(ll), _dptr(ess),
            res.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return ll.value, ess, res

    def log_likelihood(self):
        return self.results()[0]

    def estimate_moments(self, k):
        """Selection estimator of x^1..x^k (SMCEstimator at step 0,
        algorithm/smc.hpp:89-118); k = 0 removes it.  Call before iterate()."""
        check(lib.mckl_hip_smc_estimate_moments(self._h, int(k)))
        self._k = int(k)

    def estimates(self):
        """(T, k) array: row t holds sum_i W_i x_i^j, j = 1..k, of step t."""
        k = getattr(self, "_k", 0)
        out = np.zeros((self.T, max(k, 1)), np.float64)
        check(lib.mckl_hip_smc_estimates(self._h, _dptr(out)))
        return out[:, :k]

    # the library's destroy function, bound while the module is alive (at
    # interpreter exit the module's globals may already be cleared)
    _destroy = staticmethod(lib.mckl_hip_smc_destroy)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
            self._h = None


def smc_loglik(model, theta, y, N, seed=1, threshold=0.5, scheme=SYSTEMATIC):
    s = SMCSampler(model, theta, y, N, seed, threshold, scheme)
    s.iterate()
    return s.log_likelihood()


class SMCBatch:
    """Many independent filters in one persistent launch."""

    def __init__(self, model, N, theta, y, seeds, resample_threshold=0.5,
                 scheme=STRATIFIED):
        self._N = np.ascontiguousarray(N, np.uint32)
        self._th = np.ascontiguousarray(theta, np.float64).reshape(-1)
        self._y = np.ascontiguousarray(y, np.float64)
        self._seeds = np.ascontiguousarray(seeds, np.uint64)
        self.nfilters = len(self._N)
        self.T = self._y.shape[-1]
        h = ctypes.c_void_p()
        _init_device()
        check(lib.mckl_hip_smc_batch_create(
            model, self.nfilters,
            self._N.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
            _dptr(self._th), self.T, _dptr(self._y.reshape(-1)),
            self._seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            float(resample_threshold), scheme, ctypes.byref(h)))
        self._h = h

    def set_tune(self, block=256, waves_per_eu=0, split_tiles=0,
                 small_block=256):
        """Launch configuration (changes no result bit): workgroup width,
        occupancy cap, and with split_tiles > 0 a second concurrent launch
        of small_block-thread workgroups for the filters of fewer tiles."""
        t = _lib.Tune(block, waves_per_eu, split_tiles, small_block)
        check(lib.mckl_hip_smc_batch_set_tune(self._h, ctypes.byref(t)))

    def autotune(self, cache=None):
        """The library's autotuner: times its candidate configurations,
        checks they give identical bits, keeps the fastest; with `cache` (a
        file path) reuses or records the choice for this workload.  Returns
        (dict of the configuration, ms per run)."""
        t = _lib.Tune()
        ms = ctypes.c_double()
        path = None if cache is None else os.fsencode(cache)
        check(lib.mckl_hip_smc_batch_autotune(self._h, path, ctypes.byref(t),
                                              ctypes.byref(ms), _stream()))
        return ({"block": t.block, "waves_per_eu": t.waves_per_eu,
                 "split_tiles": t.split_tiles, "small_block": t.small_block},
                ms.value)

    def run(self, use_graph=True):
        check(lib.mckl_hip_smc_batch_run(self._h, int(bool(use_graph)),
                                         _stream()))

    def log_likelihood(self):
        out = np.zeros(self.nfilters, np.float64)
        check(lib.mckl_hip_smc_batch_result(self._h, _dptr(out)))
        return out

    def set_params(self, theta, seeds):
        """New theta (nfilters x 3) and seeds for the next run."""
        th = np.ascontiguousarray(theta, np.float64).reshape(-1)
        sd = np.ascontiguousarray(seeds, np.uint64)
        if th.size != 3 * self.nfilters or sd.size != self.nfilters:
            raise MCKLError("set_params: one theta row and seed per filter")
        check(lib.mckl_hip_smc_batch_set_params(
            self._h, _dptr(th),
            sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))

    # the library's destroy function, bound while the module is alive (at
    # interpreter exit the module's globals may already be cleared)
    _destroy = staticmethod(lib.mckl_hip_smc_batch_destroy)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
            self._h = None


class SMCGradient:
    """Gradient of the SMC log-likelihood estimate by the hand-written
    adjoint, for a batch of parameter points.

    theta has shape (nbatch, 3); y has shape (T,) (shared) or (nbatch, T);
    seeds one per point.  With tape=True, run() records the ancestors and
    decisions so that replay(theta) evaluates the differentiated function at
    another theta with them held fixed (central finite differences of
    replay() are what the adjoint's gradient must reproduce).
    """

    def __init__(self, model, nbatch, N, y, seeds, resample_threshold=0.5,
                 scheme=SYSTEMATIC, K=50, tape=False):
        self.nbatch = int(nbatch)
        self._y = np.ascontiguousarray(y, np.float64)
        per_point = self._y.ndim == 2
        self.T = self._y.shape[-1]
        self._seeds = np.ascontiguousarray(
            np.broadcast_to(np.asarray(seeds, np.uint64), (self.nbatch,)))
        h = ctypes.c_void_p()
        _init_device()
        check(lib.mckl_hip_smc_grad_create(
            model, self.nbatch, int(N), self.T, _dptr(self._y.reshape(-1)),
            int(per_point),
            self._seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
            float(resample_threshold), scheme, int(K), int(bool(tape)),
            ctypes.byref(h)))
        self._h = h

    def _theta(self, theta):
        th = np.ascontiguousarray(theta, np.float64).reshape(-1, 3)
        if th.shape[0] != self.nbatch:
            th = np.ascontiguousarray(np.broadcast_to(th, (self.nbatch, 3)))
        return th

    def run(self, theta):
        """Returns (loglik[nbatch], grad[nbatch, 3])."""
        th = self._theta(theta)
        ll = np.zeros(self.nbatch, np.float64)
        g = np.zeros((self.nbatch, 3), np.float64)
        check(lib.mckl_hip_smc_grad_run(self._h, _dptr(th), _dptr(ll),
                                        _dptr(g), _stream()))
        return ll, g

    def replay(self, theta):
        th = self._theta(theta)
        v = np.zeros(self.nbatch, np.float64)
        check(lib.mckl_hip_smc_grad_replay(self._h, _dptr(th), _dptr(v),
                                           _stream()))
        return v

    # the library's destroy function, bound while the module is alive (at
    # interpreter exit the module's globals may already be cleared)
    _destroy = staticmethod(lib.mckl_hip_smc_grad_destroy)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._destroy(h)
            self._h = None


def smc_loglik_grad(model, theta, y, N, seeds, resample_threshold=0.5,
                    scheme=SYSTEMATIC, K=50):
    """One-shot batched log-likelihood and gradient (no tape)."""
    th = np.ascontiguousarray(theta, np.float64).reshape(-1, 3)
    nb = th.shape[0]
    yy = np.ascontiguousarray(y, np.float64)
    sd = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, np.uint64),
                                              (nb,)))
    ll = np.zeros(nb, np.float64)
    g = np.zeros((nb, 3), np.float64)
    check(lib.mckl_hip_smc_loglik_grad(
        model, nb, _dptr(th), int(N), yy.shape[-1], _dptr(yy.reshape(-1)),
        int(yy.ndim == 2), sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
        float(resample_threshold), scheme, int(K), _dptr(ll), _dptr(g),
        _stream()))
    return ll, g


class SeedGenerator:
    """SeedGenerator<key_type of an engine> (random/seed.hpp:54-341): the
    keys RNGSetVector::reset() hands to its engines (rng_set.hpp:81-120).
    `keys(n)` returns the next n keys as a (n, 4) uint64 device tensor in
    engine-key layout and advances the internal seed by n, as n get() calls
    would."""

    def __init__(self, engine_cls=None, randomize=True):
        self.kind = (engine_cls or Philox4x32).KIND
        self.randomize = bool(randomize)
        self.seed = 1
        self.np = 1
        self.rank = 0

    def set(self, s):
        self.seed = int(s)

    def partition(self, np_, rank):
        if np_ < 1 or rank >= np_:
            raise MCKLError("partition needs np > 0 and rank < np")
        self.np, self.rank = int(np_), int(rank)

    def keys(self, n):
        import torch
        out = torch.empty((n, 4), dtype=torch.int64, device="cuda")
        check(lib.mckl_hip_seed_keys(self.kind, self.seed, self.np, self.rank,
                                     int(self.randomize), n, _ptr(out),
                                     _stream()))
        self.seed += n
        return out


class RNGSetVector:
    """RNGSetVector<RNG>(n) (random/rng_set.hpp:81-120) on the device: n
    engines keyed by the SeedGenerator, counters 0.  rand/u01 give each
    engine's first `per` outputs (engine e in rows e of the output)."""

    def __init__(self, engine_cls, n, seed_generator=None):
        if engine_cls.RBITS != 32:
            raise MCKLError("RNG sets are provided for 32-bit engines")
        self.engine_cls = engine_cls
        self.n = int(n)
        sg = seed_generator or SeedGenerator(engine_cls)
        self.keys = sg.keys(self.n)

    def engine(self, e):
        """The host engine of member e (for further calls of its own)."""
        k = [int(v) & 0xFFFFFFFFFFFFFFFF for v in self.keys[e].cpu().tolist()]
        return self.engine_cls(key=k[:self.engine_cls.KEY_WORDS])

    def rand(self, per, out):
        p = _dev(out, self.n * per, 4, "RNG set output", False)
        check(lib.mckl_hip_rngset_rand(self.engine_cls.KIND, self.n,
                                       _ptr(self.keys), per, p, _stream()))
        return out

    def u01(self, per, out, interval=_lib.U01_OO):
        import torch
        fn = (lib.mckl_hip_rngset_u01_f64 if out.dtype == torch.float64
              else lib.mckl_hip_rngset_u01_f32)
        p = _dev(out, self.n * per, 8 if out.dtype == torch.float64 else 4,
                 "RNG set output", True)
        check(fn(self.engine_cls.KIND, interval, self.n, _ptr(self.keys),
                 per, p, _stream()))
        return out


from . import pmcmc  # noqa: E402,F401  (PMCMCMutation, PMCMCBatch)
from .pmcmc import (PMCMCMutation, PMCMCBatch,  # noqa: E402,F401
                    GenericPMCMCMutation)
from .sampler import (Particle, ResampleEval, SMCEstimator,  # noqa: E402,F401
                      GenericSMCSampler, PMCMCParticle)
