"""ctypes binding of the oracle (test infrastructure only).

``O``   -- libmckl_oracle.so, the CPU restatement (mckl_oracle.c)
``REF`` -- oracle/_ref/libmckl_ref.so, the reference's own headers behind
           ref_driver.cpp, when it has been built (None otherwise)
``REF_SMC`` -- oracle/_ref/libmckl_ref_smc.so, the reference's SMCSampler
           behind ref_smc.cpp (lazy binding: its BLAS symbol is never
           called), when built
``REF_V4`` -- oracle/_ref/libmckl_ref_v4.so, the same driver as ``REF``
           built -O3 for x86-64-v4 with AES-NI (the reference's AVX-512 and
           AES-NI paths), loaded only on a host that has those features;
           timing baselines only

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "libmckl_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmckl_ref.so")
REF_SMC_PATH = os.path.join(HERE, "_ref", "libmckl_ref_smc.so")
REF_V4_PATH = os.path.join(HERE, "_ref", "libmckl_ref_v4.so")
V4_FLAGS = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl", "avx2",
            "bmi2", "fma", "aes")

PHILOX4X32, THREEFRY4X64, ARS = 1, 2, 3
CC, CO, OC, OO, CANONICAL = 0, 1, 2, 3, 4
MULTINOMIAL, STRATIFIED, SYSTEMATIC = 0, 1, 2
RESIDUAL, RESIDUAL_STRATIFIED, RESIDUAL_SYSTEMATIC = 3, 4, 5
MODEL_LG, MODEL_SV = 1, 2

O = ctypes.CDLL(ORACLE_PATH)
REF = ctypes.CDLL(REF_PATH) if os.path.exists(REF_PATH) else None
REF_SMC = (ctypes.CDLL(REF_SMC_PATH, mode=os.RTLD_LAZY)
           if os.path.exists(REF_SMC_PATH) else None)


def cpu_flags():
    """The feature flags of this host's first CPU (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


REF_V4 = (ctypes.CDLL(REF_V4_PATH)
          if os.path.exists(REF_V4_PATH) and
          all(f in cpu_flags() for f in V4_FLAGS) else None)

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t


class Eng(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("rbits", ctypes.c_int),
                ("ctr", ctypes.c_uint64 * 4), ("key", ctypes.c_uint64 * 4),
                ("buf", ctypes.c_uint8 * 32), ("index", ctypes.c_uint)]


def P(a):
    return a.ctypes.data_as(_vp)


O.mo_smc_normal.restype = ctypes.c_float
O.mo_smc_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]
O.mo_smc_run.restype = ctypes.c_double
O.mo_smc_run.argtypes = [ctypes.c_int, _vp, _vp, _sz, _sz, ctypes.c_uint64,
                         ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp, _vp]
O.mo_smc_run_est.restype = ctypes.c_double
O.mo_smc_run_est.argtypes = [ctypes.c_int, _vp, _vp, _sz, _sz, ctypes.c_uint64,
                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _vp,
                             _vp, ctypes.c_int, _vp]
O.mo_next_u32.restype = ctypes.c_uint32
O.mo_u01_u32_f64.restype = ctypes.c_double
O.mo_u01_u32_f32.restype = ctypes.c_float
O.mo_u01_canonical_f64.restype = ctypes.c_double
O.mo_u01_canonical_f32.restype = ctypes.c_float
for _lib in (REF, REF_V4):
    if _lib is not None:
        _lib.ref_trans_residual.restype = _sz
        _lib.ref_time_philox_u01.restype = ctypes.c_double
        _lib.ref_time_philox_u01_mt.restype = ctypes.c_double
        _lib.ref_time_normal_streams.restype = ctypes.c_double
        _lib.ref_time_normal_streams.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, _sz, ctypes.c_double,
            ctypes.c_double, ctypes.c_int, _vp]


def engine(kind, key, rbits=32):
    e = Eng()
    k = (ctypes.c_uint64 * 4)(*(list(key) + [0] * (4 - len(key))))
    O.mo_engine_init(ctypes.byref(e), kind, rbits, k)
    return e


def state(e):
    n = 4 if e.kind == THREEFRY4X64 else 2
    return list(e.ctr[:n]), int(e.index)


def rand_u32(e, n):
    out = np.zeros(max(n, 1), np.uint32)
    O.mo_rand_u32(ctypes.byref(e), _sz(n), P(out))
    return out[:n]


def rand_u64(e, n):
    out = np.zeros(max(n, 1), np.uint64)
    O.mo_rand_u64(ctypes.byref(e), _sz(n), P(out))
    return out[:n]


def rand_u01(e, interval, n, dtype):
    out = np.zeros(max(n, 1), dtype)
    fn = O.mo_rand_u01_f64 if np.dtype(dtype) == np.float64 else O.mo_rand_u01_f32
    fn(ctypes.byref(e), interval, _sz(n), P(out))
    return out[:n]


def rand_uniform_real(e, n, a, b, dtype):
    out = np.zeros(max(n, 1), dtype)
    if np.dtype(dtype) == np.float64:
        O.mo_rand_uniform_real_f64(ctypes.byref(e), _sz(n), ctypes.c_double(a),
                                   ctypes.c_double(b), P(out))
    else:
        O.mo_rand_uniform_real_f32(ctypes.byref(e), _sz(n), ctypes.c_float(a),
                                   ctypes.c_float(b), P(out))
    return out[:n]


def ref_uniform_real(kind, key, dtype, n, a, b, pre=0, rbits=32):
    out = np.zeros(n + 1, dtype)
    k = (ctypes.c_uint64 * 4)(*(list(key) + [0] * (4 - len(key))))
    ctr = (ctypes.c_uint64 * 4)()
    idx = ctypes.c_uint()
    REF.ref_uniform_real(kind, rbits, k, int(np.dtype(dtype) == np.float64),
                         _sz(pre),
                         _sz(n), ctypes.c_double(a), ctypes.c_double(b), P(out),
                         ctr, ctypes.byref(idx))
    nc = 4 if kind == THREEFRY4X64 else 2
    return out[:n], list(ctr[:nc]), idx.value


def rand_normal(e, n, mean, sd, dtype):
    out = np.zeros(max(n, 1), dtype)
    if np.dtype(dtype) == np.float64:
        O.mo_rand_normal_f64(ctypes.byref(e), _sz(n), ctypes.c_double(mean),
                             ctypes.c_double(sd), P(out))
    else:
        O.mo_rand_normal_f32(ctypes.byref(e), _sz(n), ctypes.c_float(mean),
                             ctypes.c_float(sd), P(out))
    return out[:n]


def rand_normal_cr(e, n, mean, sd, dtype):
    """rand_normal plus, per output, whether every libm result it depends
    on (log, and sin or cos) is the correctly rounded value."""
    out = np.zeros(max(n, 1), dtype)
    clean = np.zeros(max(n, 1), np.uint8)
    if np.dtype(dtype) == np.float64:
        O.mo_rand_normal_f64_cr(ctypes.byref(e), _sz(n), ctypes.c_double(mean),
                                ctypes.c_double(sd), P(out), P(clean))
    else:
        O.mo_rand_normal_f32_cr(ctypes.byref(e), _sz(n), ctypes.c_float(mean),
                                ctypes.c_float(sd), P(out), P(clean))
    return out[:n], clean[:n].astype(bool)


CR_LOG, CR_SIN, CR_COS, CR_LOG1P = 0, 1, 2, 3


def quad_eval(fn, x):
    """log / sin / cos of doubles in quad precision, rounded once."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(max(len(x), 1))
    O.mo_quad_eval(fn, _sz(len(x)), P(x), P(y))
    return y[:len(x)]


def quad_dd_relerr(fn, x0, hi, lo):
    x0, hi, lo = (np.ascontiguousarray(a, np.float64) for a in (x0, hi, lo))
    err = np.zeros(max(len(x0), 1))
    O.mo_quad_dd_relerr(fn, _sz(len(x0)), P(x0), P(hi), P(lo), P(err))
    return err[:len(x0)]


def philox_block(key, ctr):
    k = (ctypes.c_uint32 * 2)(*key)
    c = (ctypes.c_uint64 * 2)(*ctr)
    o = (ctypes.c_uint32 * 4)()
    O.mo_philox4x32_block(k, c, o)
    return list(o)


def threefry_block(key, ctr):
    k = (ctypes.c_uint64 * 4)(*key)
    c = (ctypes.c_uint64 * 4)(*ctr)
    o = (ctypes.c_uint64 * 4)()
    O.mo_threefry4x64_block(k, c, o)
    return list(o)


def ars_block(key, ctr):
    k = (ctypes.c_uint32 * 4)(*key)
    c = (ctypes.c_uint64 * 2)(*ctr)
    o = (ctypes.c_uint32 * 4)()
    O.mo_ars_block(k, c, o)
    return list(o)


def aes128(key, block):
    k = (ctypes.c_uint8 * 16)(*key)
    b = (ctypes.c_uint8 * 16)(*block)
    o = (ctypes.c_uint8 * 16)()
    O.mo_aes128_encrypt(k, b, o)
    return bytes(o)


def weight_set_log(v):
    v = np.ascontiguousarray(v, np.float64)
    w = np.zeros_like(v)
    ess = ctypes.c_double()
    O.mo_weight_set_log(_sz(len(v)), P(v), P(w), ctypes.byref(ess))
    return w, ess.value


def weight_add_log(w, v):
    v = np.ascontiguousarray(v, np.float64)
    w = np.array(w, np.float64)
    ess = ctypes.c_double()
    lnc = ctypes.c_double()
    O.mo_weight_add_log(_sz(len(v)), P(v), P(w), ctypes.byref(ess),
                        ctypes.byref(lnc))
    return w, ess.value, lnc.value


def u01_sequence(scheme, e, n):
    out = np.zeros(max(n, 1), np.float64)
    fn = {MULTINOMIAL: O.mo_u01_rand_sorted, STRATIFIED: O.mo_u01_rand_stratified,
          SYSTEMATIC: O.mo_u01_rand_systematic}[scheme % 3]
    fn(ctypes.byref(e), _sz(n), P(out))
    return out[:n]


def trans_u01_rep(w, u, N=None, M=None, exact_prefix=False):
    """resample_trans_u01_rep: sequential FP64 cumulative weights (the
    reference's), or with exact_prefix the exact prefix sums rounded once
    (the device's definition)."""
    w = np.ascontiguousarray(w, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    N = len(w) if N is None else N
    M = len(u) if M is None else M
    rep = np.zeros(max(N, 1), np.uint64)
    fn = (O.mo_resample_trans_u01_rep_cr if exact_prefix
          else O.mo_resample_trans_u01_rep)
    fn(_sz(N), _sz(M), P(w), P(u), P(rep))
    return rep[:N]


def rep_index(rep, M):
    rep = np.ascontiguousarray(rep, np.uint64)
    idx = np.zeros(max(M, 1), np.uint64)
    O.mo_resample_trans_rep_index(_sz(len(rep)), _sz(M), P(rep), P(idx))
    return idx[:M]


def resample(scheme, e, w, M=None):
    w = np.ascontiguousarray(w, np.float64)
    N = len(w)
    M = N if M is None else M
    rep = np.zeros(max(N, 1), np.uint64)
    idx = np.zeros(max(M, 1), np.uint64)
    O.mo_resample(scheme, _sz(N), _sz(M), ctypes.byref(e), P(w), P(rep), P(idx))
    return rep[:N], idx[:M]


def engine_text(e):
    buf = ctypes.create_string_buffer(1024)
    n = O.mo_engine_text(ctypes.byref(e), buf, _sz(1024))
    assert n > 0
    return buf.value.decode()


def seed_keys(kind, seed, n, np_=1, rank=0, randomize=True):
    """SeedGenerator<key_type>::get() n times from internal seed `seed`:
    (n, 4) uint64 keys in engine-key layout."""
    out = np.zeros((n, 4), np.uint64)
    O.mo_seed_keys(kind, ctypes.c_uint64(seed), ctypes.c_uint64(np_),
                   ctypes.c_uint64(rank), int(bool(randomize)), _sz(n), P(out))
    return out


def smc_normal(seed, i, t):
    return O.mo_smc_normal(seed, i, t)


def smc_run(model, theta, y, N, seed, threshold=0.5, scheme=SYSTEMATIC,
            nthreads=1, history=False, moments=0):
    """loglik (and with history=True the ESS and resampling flags per step;
    with moments=k > 0 also the (T, k) selection estimates of x^1..x^k)."""
    th = np.ascontiguousarray(theta, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    # the growth model (3) takes the T observations then the T drifts
    T = len(y) // 2 if model == 3 else len(y)
    ess = np.zeros(T, np.float64)
    res = np.zeros(T, np.int32)
    est = np.zeros((T, max(moments, 1)), np.float64)
    ll = O.mo_smc_run_est(model, P(th), P(y), _sz(T), _sz(N), seed, threshold,
                          scheme, nthreads, P(ess), P(res), moments, P(est))
    if moments:
        return ll, ess, res, est[:, :moments]
    if history:
        return ll, ess, res
    return ll


O.mo_smc_resample_dump.restype = _sz
O.mo_smc_resample_dump.argtypes = [ctypes.c_int, _vp, _vp, _sz, _sz,
                                   ctypes.c_uint64, ctypes.c_double,
                                   ctypes.c_int, _sz, _vp, _vp, _vp, _vp]


def smc_resample_dump(model, theta, y, N, seed, threshold, scheme, t_first=1):
    """The fused sampler's first resampling step at or after t_first (any
    scheme): (t, u01, W, u, rep) with the step's uniform draws u01
    (systematic forms: one), the normalised weights W, the uniform sequence
    u (residual schemes: the R uniforms of the remainder, R = N - sum of
    floor(N W_i)) and the replication numbers; None when no step from
    t_first on resamples."""
    th = np.ascontiguousarray(theta, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    T = len(y) // 2 if model == 3 else len(y)
    u01 = np.zeros(N, np.float64)
    W = np.zeros(N, np.float64)
    u = np.zeros(N, np.float64)
    rep = np.zeros(N, np.uint64)
    t = O.mo_smc_resample_dump(model, P(th), P(y), _sz(T), _sz(N), seed,
                               threshold, scheme, _sz(t_first), P(u01), P(W),
                               P(u), P(rep))
    if t == 0:
        return None
    n = N
    if scheme >= RESIDUAL:
        n = N - int(np.floor(float(N) * W).sum())
    if scheme in (SYSTEMATIC, RESIDUAL_SYSTEMATIC):
        return t, u01[:1], W, u[:n], rep
    return t, u01[:n], W, u[:n], rep


# ---- reference (compiled from the reference headers), when available ----
def ref_u01_trans(scheme, u01, N):
    """u01_trans_sorted / stratified / systematic of given uniforms (scheme
    0, 1, 2; systematic reads u01[0])."""
    u01 = np.ascontiguousarray(u01, np.float64)
    r = np.zeros(max(N, 1), np.float64)
    REF.ref_u01_trans(scheme, _sz(N), P(u01), P(r))
    return r[:N]


def ref_trans_residual(w, M=None):
    """resample_trans_residual: (normalised residuals, integral parts, R)."""
    w = np.ascontiguousarray(w, np.float64)
    N = len(w)
    M = N if M is None else M
    resid = np.zeros(max(N, 1), np.float64)
    integ = np.zeros(max(N, 1), np.uint64)
    R = REF.ref_trans_residual(_sz(N), _sz(M), P(w), P(resid), P(integ))
    return resid[:N], integ[:N], int(R)


def ref_trans_u01_rep(w, u, M=None):
    """resample_trans_u01_rep of normalised weights w and ordered uniforms u
    (the reference's sequential FP64 cumulative sum)."""
    w = np.ascontiguousarray(w, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    N = len(w)
    M = len(u) if M is None else M
    rep = np.zeros(max(N, 1), np.uint64)
    REF.ref_trans_u01_rep(_sz(N), _sz(M), P(w), P(u), P(rep))
    return rep[:N]


def ref_engine_ops(kind, rbits, key, sizes, ctr0=None):
    n = sum(s if s > 0 else (1 if s == 0 else 0) for s in sizes)
    out = np.zeros(max(n, 1), np.uint32 if rbits == 32 else np.uint64)
    k = (ctypes.c_uint64 * 4)(*(list(key) + [0] * (4 - len(key))))
    c0 = None if ctr0 is None else (ctypes.c_uint64 * 4)(
        *(list(ctr0) + [0] * (4 - len(ctr0))))
    sz = (ctypes.c_longlong * len(sizes))(*sizes)
    ctr = (ctypes.c_uint64 * 4)()
    idx = ctypes.c_uint()
    REF.ref_engine_ops(kind, rbits, k, c0, len(sizes), sz, P(out), ctr,
                       ctypes.byref(idx))
    nc = 4 if kind == THREEFRY4X64 else 2
    return out[:n], list(ctr[:nc]), idx.value


def ref_u01(kind, key, interval, dtype, n, pre=0, rbits=32):
    out = np.zeros(n + 1, dtype)
    k = (ctypes.c_uint64 * 4)(*(list(key) + [0] * (4 - len(key))))
    ctr = (ctypes.c_uint64 * 4)()
    idx = ctypes.c_uint()
    REF.ref_u01(kind, rbits, k, interval, int(np.dtype(dtype) == np.float64),
                _sz(pre),
                _sz(n), P(out), ctr, ctypes.byref(idx))
    nc = 4 if kind == THREEFRY4X64 else 2
    return out[:n], list(ctr[:nc]), idx.value


_UBITS = {16: np.uint16, 32: np.uint32, 64: np.uint64}


def rand_uniform_bits(e, ubits, n):
    out = np.zeros(max(n, 1), _UBITS[ubits])
    O.mo_rand_uniform_bits(ctypes.byref(e), ubits, _sz(n), P(out))
    return out[:n]


def ref_uniform_bits(kind, key, ubits, n, pre=0, rbits=32):
    out = np.zeros(n + 1, _UBITS[ubits])
    k = (ctypes.c_uint64 * 4)(*(list(key) + [0] * (4 - len(key))))
    ctr = (ctypes.c_uint64 * 4)()
    idx = ctypes.c_uint()
    REF.ref_uniform_bits(kind, rbits, k, ubits, _sz(pre), _sz(n), P(out), ctr,
                         ctypes.byref(idx))
    nc = 4 if kind == THREEFRY4X64 else 2
    return out[:n], list(ctr[:nc]), idx.value


def ref_normal(kind, key, dtype, n, mean, sd, pre=0, rbits=32):
    out = np.zeros(n + 1, dtype)
    k = (ctypes.c_uint64 * 4)(*(list(key) + [0] * (4 - len(key))))
    ctr = (ctypes.c_uint64 * 4)()
    idx = ctypes.c_uint()
    REF.ref_normal(kind, rbits, k, int(np.dtype(dtype) == np.float64), _sz(pre),
                   _sz(n),
                   ctypes.c_double(mean), ctypes.c_double(sd), P(out), ctr,
                   ctypes.byref(idx))
    nc = 4 if kind == THREEFRY4X64 else 2
    return out[:n], list(ctr[:nc]), idx.value


def ref_resample(scheme, kind, key, w, M=None):
    w = np.ascontiguousarray(w, np.float64)
    N = len(w)
    M = N if M is None else M
    rep = np.zeros(max(N, 1), np.uint64)
    idx = np.zeros(max(M, 1), np.uint64)
    k = (ctypes.c_uint64 * 4)(*(list(key) + [0] * (4 - len(key))))
    REF.ref_resample(scheme, _sz(N), _sz(M), kind, k, P(w), P(rep), P(idx))
    return rep[:N], idx[:M]


def ref_weight(mode, v, w=None, ess=0.0):
    v = np.ascontiguousarray(v, np.float64)
    w = np.zeros_like(v) if w is None else np.array(w, np.float64)
    e = ctypes.c_double(ess)
    REF.ref_weight(mode, _sz(len(v)), P(v), P(w), ctypes.byref(e))
    return w, e.value


def ref_u01_sequence(scheme, kind, key, n):
    out = np.zeros(max(n, 1), np.float64)
    k = (ctypes.c_uint64 * 4)(*(list(key) + [0] * (4 - len(key))))
    ctr = (ctypes.c_uint64 * 4)()
    idx = ctypes.c_uint()
    REF.ref_u01_sequence(scheme, kind, k, _sz(n), P(out), ctr, ctypes.byref(idx))
    return out[:n]


class ResampleIndex:
    """Restatement of ResampleIndex<IntType> (algorithm/resample.hpp:597-832)
    on numpy uint32 vectors, including its bookkeeping: reset() keeps the
    vectors (:628), clear() frees them and keeps num_iter (:631), push_back
    writes vector num_iter-1 when it exists and appends otherwise
    (:634-656), index() starts from the last vector held whatever iter_back
    is (:675).  Queries the reference guards with runtime_assert raise
    ValueError; reading a vector clear() has freed (undefined in the
    reference) raises IndexError.  An entry that does not index the vector
    it is read against traces to 0xFFFFFFFF (undefined in the reference)."""

    BAD = 0xFFFFFFFF

    def __init__(self):
        self.num_iter = 0
        self.vec = []

    def size(self, iter=None):
        if iter is None:                                   # :609-615
            return len(self.vec[-1]) if self.vec else 0
        return len(self.vec[iter]) if iter < len(self.vec) else 0  # :618-625

    def reset(self):
        self.num_iter = 0

    def clear(self):
        self.vec = []

    def push_back(self, N, idx=None):
        self.num_iter += 1
        v = np.arange(N, dtype=np.uint32) if idx is None else \
            np.array(idx[:N], dtype=np.uint32)
        if len(self.vec) < self.num_iter:
            self.vec.append(v)
        else:
            self.vec[self.num_iter - 1] = v

    @staticmethod
    def _gather(v, a):
        a = np.asarray(a, dtype=np.uint64)
        ok = a < len(v)
        out = np.full(a.shape, ResampleIndex.BAD, dtype=np.uint32)
        out[ok] = v[a[ok]]
        return out

    def _check(self, iter_back, iter):
        if not (iter <= iter_back < self.num_iter):
            raise ValueError("iteration numbers out of range")

    def index(self, ids, iter_back=None, iter=0):
        """index(id, iter_back, iter) (:669-681) for an array of ids."""
        iter_back = self.num_iter - 1 if iter_back is None else iter_back
        self._check(iter_back, iter)
        if not self.vec or (iter_back > iter and iter_back - 1 >= len(self.vec)):
            raise IndexError("vector released by clear()")
        a = self._gather(self.vec[-1], ids)
        while iter_back > iter:
            iter_back -= 1
            a = self._gather(self.vec[iter_back], a)
        return a

    def shape(self, iter_back=None, iter=0):
        iter_back = self.num_iter - 1 if iter_back is None else iter_back
        self._check(iter_back, iter)
        if iter_back >= len(self.vec):
            raise IndexError("vector released by clear()")
        return len(self.vec[iter_back]), iter_back - iter + 1   # :683-715

    def index_matrix(self, layout, iter_back=None, iter=0):
        """read_index_matrix (:741-810): column R-1 is vector iter_back, column
        c is vector iter_back-(R-1-c) gathered at column c+1; returned flat
        in the requested layout (101 row-major, 102 column-major)."""
        iter_back = self.num_iter - 1 if iter_back is None else iter_back
        N, R = self.shape(iter_back, iter)
        cols = [None] * R
        cols[R - 1] = self.vec[iter_back].copy()
        for r in range(1, R):
            cols[R - 1 - r] = self._gather(self.vec[iter_back - r],
                                           cols[R - r])
        m = np.stack(cols, axis=1) if R else np.zeros((N, 0), np.uint32)
        return (m if layout == 101 else m.T).reshape(-1)


def ref_resample_index(script):
    """Run a ResampleIndex script through the compiled reference.  script:
    list of ("push", N, idx-or-None) / ("reset",) / ("clear",) /
    ("index", ids, iter_back, iter) / ("matrix", layout, iter_back, iter) /
    ("size", iter-or-None) / ("num_iter",).  Returns the query results."""
    out = []
    REF.ref_ri_new()
    for op in script:
        if op[0] == "push":
            N, idx = op[1], op[2]
            if idx is None:
                REF.ref_ri_push(_sz(N), None)
            else:
                a = np.ascontiguousarray(idx[:N], np.uint32)
                REF.ref_ri_push(_sz(N), P(a))
        elif op[0] == "reset":
            REF.ref_ri_reset()
        elif op[0] == "clear":
            REF.ref_ri_clear()
        elif op[0] == "num_iter":
            out.append(REF.ref_ri_size(ctypes.c_int(-2), _sz(0)))
        elif op[0] == "size":
            it = op[1]
            out.append(REF.ref_ri_size(ctypes.c_int(-1 if it is None else 0),
                                       _sz(0 if it is None else it)))
        elif op[0] == "index":
            ids = np.ascontiguousarray(op[1], np.uint32)
            r = np.zeros(max(len(ids), 1), np.uint32)
            REF.ref_ri_index(_sz(len(ids)), P(ids), _sz(op[2]), _sz(op[3]),
                             P(r))
            out.append(r[:len(ids)])
        elif op[0] == "matrix":
            nr = REF.ref_ri_nrow(_sz(op[2]))
            nc = op[2] - op[3] + 1
            r = np.zeros(max(nr * nc, 1), np.uint32)
            REF.ref_ri_matrix(op[1], _sz(op[2]), _sz(op[3]), P(r))
            out.append(r[:nr * nc])
    return out


if REF is not None:
    REF.ref_ri_size.restype = ctypes.c_size_t
    REF.ref_ri_nrow.restype = ctypes.c_size_t


def pmmh(model, y, N, M, prior, updates, params, iters, seed=1,
         threshold=0.5, scheme=STRATIFIED):
    """Restatement of PMCMCMutation::operator() (algorithm/pmcmc.hpp:125-158)
    for C chains side by side (C = 1 is the single mutation): params (C, 3)
    is moved in place over iterations 0..iters-1; returns the acceptance
    flags (iters, C) and the final log normalising constants.  prior maps
    (C, 3) to C log densities; updates are f(engine, P) -> C log proposal
    ratios drawing from the chains' shared Philox4x32(seed) engine as the
    device drivers do; the filter of run r, chain c is smc_run with seed
    seed + 2^32 (r C + c + 1) on y[:M].  A proposal with prior -inf is
    rejected without a run.  The acceptance draw is log U01Canonical<double>
    (random/u01_distribution.hpp:42,225-233)."""
    P0 = np.asarray(params, np.float64)
    C = P0.shape[0]
    e = engine(PHILOX4X32, [seed & 0xFFFFFFFF, 0])
    runs = [0]

    def loglik(P, ok):
        base = runs[0] * C
        runs[0] += 1
        out = np.full(C, -np.inf)
        for c in range(C):
            if ok[c]:
                s = (seed & 0xFFFFFFFF) + ((base + 1 + c) << 32)
                out[c] = smc_run(model, P[c], y[:M], N, s, threshold, scheme,
                                 nthreads=4)
        return out

    acc_hist = np.zeros((iters, C), np.int32)
    lnc = loglik(P0, np.ones(C, bool))
    for it in range(1, iters):
        prob = -np.asarray(prior(P0), np.float64) - lnc
        P = P0.copy()
        for f in updates:
            prob = prob + np.asarray(f(e, P), np.float64)
        prob = prob + np.asarray(prior(P), np.float64)
        ok = prob > -np.inf
        new = loglik(P, ok)
        prob = prob + new
        u = np.log(rand_u01(e, CANONICAL, C, np.float64))
        acc = u < prob
        P0[acc] = P[acc]
        lnc = np.where(acc, new, lnc)
        acc_hist[it] = acc
    params[...] = P0
    return acc_hist, lnc


def pmmh_random_walk(scales):
    """Oracle twin of the random-walk update: p += scales * z with z from
    rand(engine, Normal<double>(0, 1), C * 3)."""
    s = np.asarray(scales, np.float64)

    def f(e, P):
        z = rand_normal(e, P.size, 0.0, 1.0, np.float64).reshape(P.shape)
        P += s * z
        return np.zeros(P.shape[0])

    return f


# ---------------------------------------------------------------------------
# SMCSampler<T> with evaluation objects (algorithm/smc.hpp:151-421)
# ---------------------------------------------------------------------------

class Particle:
    """Particle<T> (core/particle.hpp:277-360) over the oracle's primitives:
    state x (N x dim float64, row-major), weights W with the ESS and log
    normaliser increment of the last weight operation, engine rng (the
    one ResampleEval draws from, Philox4x32 keyed by key)."""

    def __init__(self, N, dim=1, key=(1, 0)):
        self.x = np.zeros((N, dim))
        self.rng = engine(PHILOX4X32, list(key))
        self.set_equal(N)

    def size(self):
        return len(self.W)

    # Weight::set_equal (core/weight.hpp:76-80)
    def set_equal(self, n=None):
        n = self.size() if n is None else n
        self.W = np.full(n, 1.0 / n)
        self.ess = float(n)
        self.lnc = 0.0

    def set_log(self, v):
        self.W, self.ess = weight_set_log(v)
        self.lnc = 0.0

    def add_log(self, v):
        self.W, self.ess, self.lnc = weight_add_log(self.W, v)

    # Particle::select (core/particle.hpp:324-331)
    def select(self, idx):
        idx = np.asarray(idx, np.int64)
        self.x = self.x[idx].copy()
        self.set_equal(len(idx))


def resample_eval(scheme):
    """ResampleEval (algorithm/resample.hpp:470-511) of a built-in scheme:
    replication numbers from Particle::rng() and the weights, the stable
    rep -> index transform, Particle::select."""

    def ev(it, p):
        _, idx = resample(scheme, p.rng, p.W)
        p.select(idx)

    return ev


class PMCMCParticle(Particle):
    """A Particle whose state also carries PMCMCStateMatrix's parameter and
    log normalising constant (algorithm/pmcmc.hpp:46-66)."""

    def __init__(self, N, dim=1, key=(1, 0)):
        super().__init__(N, dim, key)
        self.param = None
        self._log_nc = 0.0

    def log_nc(self):
        return self._log_nc

    def add_log_nc(self, nc):
        self._log_nc += nc

    def reset(self, p, nc=0.0):
        self.param = np.array(p, np.float64)
        self._log_nc = nc


class GenericSMC:
    """SMCSampler<T>::iterate/do_iterate (algorithm/smc.hpp:352-406):
    selection evals, selection estimators, record the size and the ESS,
    resample evals when ESS < N * threshold, resample estimators, mutation
    evals, mutation estimators.  Estimators are callables est(iter, p)
    whose results are appended to self.estimates[slot]."""

    def __init__(self, N, dim=1, key=(1, 0), particle=None):
        self.particle = particle if particle is not None else Particle(
            N, dim, key)
        self.evals = [[], [], []]
        self.estimators = [[], [], []]
        self.estimates = [[], [], []]
        self.threshold = -np.inf
        self.size_history = []
        self.ess_history = []
        self.iter = 0

    def clear(self):
        """SMCSampler::clear (smc.hpp:205-212): the history only; the
        particle system is kept."""
        self.iter = 0
        self.size_history = []
        self.ess_history = []
        self.estimates = [[], [], []]

    def iterate(self, n=1):
        p = self.particle
        for _ in range(n):
            for ev in self.evals[0]:
                ev(self.iter, p)
            self._estimate(0)
            self.size_history.append(p.size())
            self.ess_history.append(p.ess)
            if p.ess < p.size() * self.threshold:
                for ev in self.evals[1]:
                    ev(self.iter, p)
            self._estimate(1)
            for ev in self.evals[2]:
                ev(self.iter, p)
            self._estimate(2)
            self.iter += 1

    def _estimate(self, slot):
        for est in self.estimators[slot]:
            self.estimates[slot].append(est(self.iter, self.particle))


def ref_smc_lg(N, T, scheme, threshold, key, a, sx, sy, mut, y, z):
    """The reference's own SMCSampler on the table-noise LG model of
    ref_smc.cpp; returns (ess history, size history, final x)."""
    y = np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    ess = np.zeros(T)
    size = np.zeros(T, np.uint64)
    x = np.zeros(N)
    k = (ctypes.c_uint64 * 2)(*[int(v) for v in key])
    REF_SMC.ref_smc_lg(_sz(N), _sz(T), scheme, ctypes.c_double(threshold), k,
                       ctypes.c_double(a), ctypes.c_double(sx),
                       ctypes.c_double(sy), ctypes.c_double(mut), P(y), P(z),
                       P(ess), P(size), P(x))
    return ess, size, x


def lg_table_model(smc, y, z, a, sx, sy, mut=1.0):
    """The table-noise LG evaluation objects of ref_smc.cpp on a GenericSMC
    (selection: move and add_log; mutation: x *= mut unless mut = 1)."""
    N = smc.particle.size()

    def sel(t, p):
        x = p.x[:, 0]
        x = z[:N].copy() if t == 0 else a * x + sx * z[t * N:(t + 1) * N]
        p.x[:, 0] = x
        d = (y[t] - x) / sy
        p.add_log(-0.5 * (d * d))

    smc.evals[0].append(sel)
    if mut != 1.0:
        def mu(t, p):
            p.x *= mut
        smc.evals[2].append(mu)


def generic_pmmh_lg(N, T, iters, scheme, threshold, key, scale, theta0, y, z):
    """PMCMCMutation<Param, T>::operator() (algorithm/pmcmc.hpp:125-158)
    over GenericSMC on the model of ref_smc.cpp's ref_pmcmc_lg: the filter's
    selection eval moves with the table noise at the state's parameter,
    adds log(sum_i W_i exp(v_i)) (in order, libm exp and log) to the log
    normalising constant, then add_log(v); the update eval is a Gaussian
    random walk from rand(rng, Normal(0, 1), 3); the acceptance draw is
    log(U01Canonical<double>) from the same engine.  Iteration 0 runs the
    filter at the starting parameter; a rejected step restores the old
    parameter and log normalising constant but keeps the particles, as the
    reference does.  Returns (accepted, params, log_nc) per iteration."""
    import math
    part = PMCMCParticle(N, 1, key)
    smc = GenericSMC(N, 1, key, particle=part)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)

    def sel(t, p):
        th = p.param
        x = p.x[:, 0]
        x = z[:N].copy() if t == 0 else th[0] * x + th[1] * z[t * N:(t + 1) * N]
        p.x[:, 0] = x
        d = (y[t] - x) / th[2]
        v = -0.5 * (d * d)
        s = 0.0
        W = p.W
        for i in range(N):
            s += float(W[i]) * math.exp(float(v[i]))
        p.add_log_nc(math.log(s))
        p.add_log(v)

    smc.evals[0].append(sel)
    smc.evals[1].append(resample_eval(scheme))
    smc.threshold = threshold
    scale = np.asarray(scale, np.float64)

    def prior(q):
        return 0.0 if abs(q[0]) < 1 and q[1] > 0 and q[2] > 0 else -np.inf

    param = np.array(theta0, np.float64)
    acc = np.zeros(iters, np.int32)
    params = np.zeros((iters, 3))
    lnc_out = np.zeros(iters)
    for it in range(iters):
        if it == 0:
            smc.clear()
            part.reset(param, 0.0)
            smc.iterate(T)
            a = 0
        else:
            lnc = part.log_nc()
            prob = -prior(param) - lnc
            q = param.copy()
            zz = rand_normal(part.rng, 3, 0.0, 1.0, np.float64)
            q = q + scale * zz
            prob += 0.0
            prob += prior(q)
            smc.clear()
            part.reset(q, 0.0)
            smc.iterate(T)
            prob += part.log_nc()
            u = math.log(float(rand_u01(part.rng, CANONICAL, 1, np.float64)[0]))
            if u < prob:
                param = q
                a = 1
            else:
                part.reset(param, lnc)
                a = 0
        acc[it] = a
        params[it] = param
        lnc_out[it] = part.log_nc()
    return acc, params, lnc_out


def ref_pmcmc_lg(N, T, iters, scheme, threshold, key, scale, theta0, y, z):
    """The reference's own PMCMCMutation on the same model (ref_smc.cpp)."""
    y = np.ascontiguousarray(y, np.float64)
    z = np.ascontiguousarray(z, np.float64)
    sc = np.ascontiguousarray(scale, np.float64)
    th = np.ascontiguousarray(theta0, np.float64)
    acc = np.zeros(iters, np.int32)
    params = np.zeros(iters * 3)
    lnc = np.zeros(iters)
    k = (ctypes.c_uint64 * 2)(*[int(v) for v in key])
    REF_SMC.ref_pmcmc_lg(_sz(N), _sz(T), _sz(iters), scheme,
                         ctypes.c_double(threshold), k, P(sc), P(th), P(y),
                         P(z), P(acc), P(params), P(lnc))
    return acc, params.reshape(iters, 3), lnc
