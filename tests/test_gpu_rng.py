"""GPU parity of the engines, raw fills, U01 and Normal fills.

The HIP path (libmckl_hip.so through its C ABI) is compared with
  * the reference's own KAT files (random_<engine>.data, 1000 words from a
    default-constructed engine, example/random/include/random_rng.hpp:80-104),
  * fixtures written from the compiled reference (tests/golden), and
  * the CPU oracle, which is itself pinned bit-exact to the reference by
    tests/test_oracle_*.py.
Integer and U01 outputs must be bit-exact.  Normal outputs: the device takes
the correctly rounded log, sin and cos, the reference glibc's, which misround
a small fraction of inputs.  Every value whose glibc results are all
correctly rounded (checked in quad precision by the oracle) must be
bit-exact; the others are counted as explained and must stay within 2 ulp.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

ENGINES = [
    ("philox4x32", orc.PHILOX4X32, 32),
    ("philox4x32_64", orc.PHILOX4X32, 64),
    ("threefry4x64", orc.THREEFRY4X64, 32),
    ("threefry4x64_64", orc.THREEFRY4X64, 64),
    ("ars", orc.ARS, 32),
    ("ars_64", orc.ARS, 64),
]


def _m():
    import mckl_amd as m
    return m


def _cls(kind, rbits):
    m = _m()
    return {(1, 32): m.Philox4x32, (1, 64): m.Philox4x32_64,
            (2, 32): m.Threefry4x64, (2, 64): m.Threefry4x64_64,
            (3, 32): m.ARS, (3, 64): m.ARS_64}[(kind, rbits)]


def _u(t):
    a = t.cpu().numpy()
    return a.view({4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def _tensor(n, itemsize, dev, fp=False):
    import torch
    if fp:
        dt = torch.float64 if itemsize == 8 else torch.float32
    else:
        dt = torch.int64 if itemsize == 8 else torch.int32
    return torch.full((max(n, 1),), -1, dtype=dt, device=dev)


def _predraw(m, rng, e, pre, rbits, dev):
    """pre raw result words (u32 or u64) from the device engine and the
    oracle's, before a distribution call"""
    if pre:
        m.rand(rng, pre, _tensor(pre, rbits // 8, dev))
        if rbits == 64:
            orc.rand_u64(e, pre)
        else:
            orc.rand_u32(e, pre)


def ulp_diff(a, b):
    if a.dtype == np.float32:
        ia = a.view(np.int32).astype(np.int64)
        ib = b.view(np.int32).astype(np.int64)
    else:
        ia = a.view(np.int64).astype(object)
        ib = b.view(np.int64).astype(object)
        d = np.abs(np.array(ia - ib, dtype=object))
        return np.array([int(x) for x in d], dtype=np.int64)
    return np.abs(ia - ib)


@pytest.mark.parametrize("name,kind,rbits", ENGINES)
def test_engine_kat(dev, name, kind, rbits):
    m = _m()
    kat = np.array([int(l) for l in open(os.path.join(
        GOLDEN, "random_%s.data" % name))], dtype=np.uint64)
    rng = _cls(kind, rbits)()   # default: seed(1), key {1, 0, ...}
    out = _tensor(len(kat), rbits // 8, dev)
    m.rand(rng, len(kat), out)
    got = _u(out)[:len(kat)].astype(np.uint64)
    assert np.array_equal(got, kat)


def test_config1_philox_stream(dev):
    m = _m()
    cfg = json.load(open(os.path.join(GOLDEN, "config1.json")))
    n = cfg["n"]
    rng = m.Philox4x32(key=cfg["key"])
    out = _tensor(n, 4, dev)
    m.rand(rng, n, out)
    a = _u(out)
    assert [int(v) for v in a[:64]] == cfg["head_u32"]
    assert [int(v) for v in a[-64:]] == cfg["tail_u32"]
    assert hashlib.sha256(a.tobytes()).hexdigest() == cfg["sha256_u32"]
    assert rng.ctr() == cfg["state_after_u32"]["ctr"]
    assert rng.index == cfg["state_after_u32"]["index"]

    rng = m.Philox4x32(key=cfg["key"])
    f = _tensor(n, 8, dev, fp=True)
    m.rand(rng, m.U01OODistribution("float64"), n, f)
    fa = f.cpu().numpy()
    assert hashlib.sha256(fa.tobytes()).hexdigest() == cfg["sha256_u01oo_f64"]
    assert rng.ctr() == cfg["state_after_u01"]["ctr"]
    assert rng.index == cfg["state_after_u01"]["index"]
    # the U01OO values of the same stream, by the closed form
    assert np.array_equal(fa, a.astype(np.float64) * 2.0 ** -32 + 2.0 ** -33)
    assert fa.min() > 0.0 and fa.max() < 1.0


SEQS = [
    [1, 1, 1, 1, 1],
    [3, 4097, 0, 5, -7, 2, 8, 13],
    [4, 8, 16, 0, 0, 0, 1000, -1000000, 3],
    [0, 7, -3, 65537, 1, 0, 31],
    [2, -1, -1, -1, 9, 100003],
]


@pytest.mark.parametrize("name,kind,rbits", ENGINES)
@pytest.mark.parametrize("seq", range(len(SEQS)))
def test_engine_call_sequences(dev, name, kind, rbits, seq):
    """Bulk calls of ragged sizes, scalar calls (0) and discards (<0):
    stream and final (ctr, index) equal the oracle's CounterEngine."""
    m = _m()
    key = [0x12345678, 0x9ABCDEF0, 3, 4][:4 if kind != 1 else 2]
    rng = _cls(kind, rbits)(key=key)
    e = orc.engine(kind, key, rbits)
    for s in SEQS[seq]:
        if s > 0:
            out = _tensor(s, rbits // 8, dev)
            m.rand(rng, s, out)
            got = _u(out)[:s]
            want = orc.rand_u32(e, s) if rbits == 32 else orc.rand_u64(e, s)
            assert np.array_equal(got, want)
        elif s < 0:
            rng.discard(-s)
            orc.O.mo_discard(ctypes.byref(e), ctypes.c_uint64(-s))
        else:
            v = rng()
            w = int(orc.O.mo_next_u32(ctypes.byref(e))) if rbits == 32 else \
                int(orc.rand_u64(e, 1)[0])
            assert v == w
        assert rng.ctr() == orc.state(e)[0]
        assert rng.index == orc.state(e)[1]


def test_counter_carry(dev):
    """Blocks across the 2^64 boundary of the low counter word."""
    m = _m()
    for kind, rbits in ((1, 32), (2, 32), (3, 32)):
        key = [5, 6, 7, 8][:2 if kind == 1 else 4]
        start = [2 ** 64 - 3, 2 ** 64 - 1, 7, 0][:4 if kind == 2 else 2]
        rng = _cls(kind, rbits)(key=key)
        rng.set_ctr(start)
        e = orc.engine(kind, key, rbits)
        c = (ctypes.c_uint64 * 4)(*(start + [0] * (4 - len(start))))
        orc.O.mo_engine_set_ctr(ctypes.byref(e), c)
        out = _tensor(101, 4, dev)
        m.rand(rng, 101, out)
        assert np.array_equal(_u(out)[:101], orc.rand_u32(e, 101))
        assert rng.ctr() == orc.state(e)[0]


@pytest.mark.parametrize("kind", [1, 2, 3])
def test_raw_kernel_abi(dev, kind):
    """mckl_hip_<engine>_kernel: n blocks for ctr+1..ctr+n, ctr += n."""
    import torch
    m = _m()
    lib = m.lib
    n = 777
    w32 = 8 if kind == 2 else 4
    out = torch.zeros(n * w32, dtype=torch.int32, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    ctr = (ctypes.c_uint64 * 4)(2 ** 64 - 100, 9, 0, 0)
    if kind == 1:
        key = (ctypes.c_uint32 * 2)(0xDEADBEEF, 0x12345678)
        m.check(lib.mckl_hip_philox4x32_kernel(ctr, n, ctypes.c_void_p(
            out.data_ptr()), key, st))
        want = []
        c = [2 ** 64 - 100, 9]
        for i in range(n):
            c = [(c[0] + 1) % 2 ** 64, c[1] + (1 if c[0] == 2 ** 64 - 1 else 0)]
            want += orc.philox_block([0xDEADBEEF, 0x12345678], c)
        end = [(2 ** 64 - 100 + n) % 2 ** 64, 10]
    elif kind == 2:
        key = (ctypes.c_uint64 * 4)(1, 2, 3, 4)
        m.check(lib.mckl_hip_threefry4x64_kernel(ctr, n, ctypes.c_void_p(
            out.data_ptr()), key, st))
        want = []
        for i in range(n):
            v = 2 ** 64 - 100 + i + 1
            blk = orc.threefry_block([1, 2, 3, 4],
                                     [v % 2 ** 64, 9 + v // 2 ** 64, 0, 0])
            for x in blk:
                want += [x & 0xFFFFFFFF, x >> 32]
        end = [(2 ** 64 - 100 + n) % 2 ** 64, 10]
    else:
        key = (ctypes.c_uint32 * 4)(11, 12, 13, 14)
        m.check(lib.mckl_hip_ars_kernel(ctr, n, ctypes.c_void_p(
            out.data_ptr()), key, st))
        want = []
        for i in range(n):
            v = 2 ** 64 - 100 + i + 1
            want += orc.ars_block([11, 12, 13, 14],
                                  [v % 2 ** 64, 9 + v // 2 ** 64])
        end = [(2 ** 64 - 100 + n) % 2 ** 64, 10]
    assert np.array_equal(_u(out), np.array(want, np.uint32))
    assert list(ctr[:2]) == end


@pytest.mark.parametrize("rbits", [32, 64])
@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("interval", [0, 1, 2, 3, 4])
def test_u01_fills(dev, kind, dtype, interval, rbits):
    """rand(rng, U01xx<T>, n, r) bit-exact, after 0..7 pre-drawn words (odd
    offsets exercise the straddling Canonical<double> path); the *_64
    engines take one u64 word per value."""
    m = _m()
    dist = [m.U01CCDistribution, m.U01CODistribution, m.U01OCDistribution,
            m.U01OODistribution, m.U01CanonicalDistribution][interval](dtype)
    isz = np.dtype(dtype).itemsize
    for pre, n in ((0, 5000), (1, 4099), (3, 2048), (2, 1), (7, 12345),
                   (5, 0)):
        key = [pre + 1, 2, 3, 4][:2 if kind == 1 else 4]
        rng = _cls(kind, rbits)(key=key)
        e = orc.engine(kind, key, rbits)
        _predraw(m, rng, e, pre, rbits, dev)
        out = _tensor(n, isz, dev, fp=True)
        m.rand(rng, dist, n, out)
        got = out.cpu().numpy()[:n]
        want = orc.rand_u01(e, interval, n, dtype)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), \
            (pre, n)
        assert rng.ctr() == orc.state(e)[0] and rng.index == orc.state(e)[1]


@pytest.mark.parametrize("rbits", [32, 64])
@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uniform_real_fills(dev, kind, dtype, rbits):
    """rand(rng, UniformRealDistribution<T>(a, b), n, r) bit-exact against
    the oracle (pinned to the compiled reference), after pre-drawn words."""
    m = _m()
    isz = np.dtype(dtype).itemsize
    for pre, n, a, b in ((0, 5000, 0.0, 1.0), (3, 2049, -2.5, 7.25),
                         (1, 1, 1e-3, 1e-3 + 1e-9), (0, 40961, -1e6, -3.0),
                         (5, 0, 0.0, 1.0), (2, 3001, 100.0, 1e8)):
        key = [pre + 3, 1, 4, 1][:2 if kind == 1 else 4]
        rng = _cls(kind, rbits)(key=key)
        e = orc.engine(kind, key, rbits)
        _predraw(m, rng, e, pre, rbits, dev)
        out = _tensor(n, isz, dev, fp=True)
        m.rand(rng, m.UniformRealDistribution(a, b, dtype), n, out)
        got = out.cpu().numpy()[:n]
        want = orc.rand_uniform_real(e, n, a, b, dtype)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), \
            (pre, n, a, b)
        assert rng.ctr() == orc.state(e)[0] and rng.index == orc.state(e)[1]
    with pytest.raises(m.MCKLError):
        m.UniformRealDistribution(1.0, 1.0)
    rng = _cls(kind, 32)(key=[1, 2, 3, 4][:2 if kind == 1 else 4])
    assert m.lib.mckl_hip_rand_uniform_real_f64(
        ctypes.byref(rng.state), 4, 2.0, 1.0, None, None) == m._lib.EINVAL


def test_misaligned_output(dev):
    """Output pointer not 16-byte aligned: element-wise stores."""
    m = _m()
    rng = m.Philox4x32(key=[1, 2])
    e = orc.engine(1, [1, 2])
    out = _tensor(5001, 4, dev)
    m.rand(rng, 5000, out[1:])
    assert np.array_equal(_u(out)[1:5001], orc.rand_u32(e, 5000))
    f = _tensor(5001, 8, dev, fp=True)
    m.rand(rng, m.U01OODistribution("float64"), 4999, f[1:])
    assert np.array_equal(f.cpu().numpy()[1:5000],
                          orc.rand_u01(e, orc.OO, 4999, np.float64))


NORMAL_CASES = [
    # (pre, n, mean, sd)
    (0, 8192, 0.0, 1.0),
    (0, 2048 * 3 + 7, 0.0, 1.0),
    (1, 5001, 2.5, 0.3),
    (3, 4097, 1.0, 2.0),
    (0, 1, 0.0, 1.0),
    (2, 3, -1.0, 0.5),
    (0, 100000, 2.5, 0.3),
]


def term_ulps(got, want, mean):
    """|got - want| in ulps of the larger of |out| and |out - mean|: with
    mean != 0, mean + sd*s*sin cancels near zero and a raw ulp count of the
    tiny result says nothing about the accuracy of the computation."""
    g = got.astype(np.float64)
    w = want.astype(np.float64)
    scale = np.maximum(np.abs(w), np.abs(w - mean)).astype(got.dtype)
    return np.abs(g - w) / np.spacing(scale).astype(np.float64)


def normal_parity(got, want, clean, mean):
    """(explained differences, histogram of their ulp distances); asserts
    that every value glibc computes from correctly rounded libm results is
    bit-exact and that the explained ones are within 2 ulp."""
    d = term_ulps(got, want, mean) if mean != 0.0 else ulp_diff(got, want)
    bad = np.nonzero((d > 0) & clean)[0]
    assert len(bad) == 0, ("unexplained differences", bad[:10],
                           got[bad[:10]], want[bad[:10]])
    expl = d[~clean]
    assert expl.size == 0 or expl.max() <= 2.0, expl.max()
    hist = np.bincount(np.ceil(expl).astype(np.int64), minlength=3)[:3]
    return int((expl > 0).sum()), hist


@pytest.mark.parametrize("rbits", [32, 64])
@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_normal_fills(dev, kind, dtype, rbits):
    m = _m()
    isz = np.dtype(dtype).itemsize
    nexpl = 0
    total = 0
    hist = np.zeros(3, np.int64)
    nclean = 0
    for pre, n, mean, sd in NORMAL_CASES:
        key = [pre + 7, 1, 2, 3][:2 if kind == 1 else 4]
        rng = _cls(kind, rbits)(key=key)
        e = orc.engine(kind, key, rbits)
        _predraw(m, rng, e, pre, rbits, dev)
        out = _tensor(n, isz, dev, fp=True)
        m.rand(rng, m.NormalDistribution(mean, sd, dtype), n, out)
        got = out.cpu().numpy()[:n]
        want, clean = orc.rand_normal_cr(e, n, mean, sd, dtype)
        k, h = normal_parity(got, want, clean, mean)
        nexpl += k
        hist += h
        nclean += int(clean.sum())
        total += n
        assert rng.ctr() == orc.state(e)[0] and rng.index == orc.state(e)[1]
    print("normal kind=%d rbits=%d %s: %d of %d values from correctly "
          "rounded libm results, all bit-exact; %d explained differences "
          "(ulp histogram 0/1/2: %s)" % (kind, rbits, np.dtype(dtype).name,
                                         nclean, total, nexpl, hist.tolist()))


@pytest.mark.parametrize("ename,kind", [("threefry", 2), ("ars", 3)])
@pytest.mark.parametrize("dname", ["f32", "f64"])
def test_config2_streams(dev, ename, kind, dname):
    """64 streams x 2^22 Normal values in one launch: windows against the
    reference fixtures (bit-exact where glibc rounds correctly, explained
    otherwise), engine states exact, moments against theory."""
    import torch
    m = _m()
    fx = np.load(os.path.join(GOLDEN, "normal_streams.npz"))
    n = 1 << 22
    dt, mean, sd = ((torch.float32, 0.0, 1.0) if dname == "f32"
                    else (torch.float64, 2.5, 0.3))
    npdt = np.float32 if dname == "f32" else np.float64
    cls = m.Threefry4x64 if kind == 2 else m.ARS
    engines = [cls(key=[s, 0, 0, 0]) for s in range(1, 65)]
    out = torch.empty(64 * n, dtype=dt, device=dev)
    m.streams_normal(engines, n, mean, sd, out)
    r = out.view(64, n)
    nexpl = 0
    hist = np.zeros(3, np.int64)
    for s in (1, 2, 64):
        p = "%s_%s_%d" % (ename, dname, s)
        row = r[s - 1].cpu().numpy()
        for part, start in (("head", 0), ("mid", n // 2 - 2048),
                            ("tail", n - 4096)):
            # windows start on chunk boundaries (2048 floats, 1024
            # doubles), so the oracle regenerates them after a discard
            e = orc.engine(kind, [s, 0, 0, 0])
            orc.O.mo_discard(ctypes.byref(e), ctypes.c_uint64(start))
            want, clean = orc.rand_normal_cr(e, 4096, mean, sd, npdt)
            assert np.array_equal(want, fx[p + "_" + part]), (p, part)
            k, h = normal_parity(row[start:start + 4096], want, clean, mean)
            nexpl += k
            hist += h
        st = fx[p + "_state"]
        e = engines[s - 1]
        assert e.ctr() == [int(v) for v in st[:-1]] and e.index == int(st[-1])
    print("config 2 %s %s windows: %d explained differences (0/1/2 ulp: %s)"
          % (ename, dname, nexpl, hist.tolist()))
    x = r.double()
    mu = x.mean().item()
    var = x.var().item()
    N = 64 * n
    assert abs(mu - mean) < 6 * sd / np.sqrt(N)
    assert abs(var - sd * sd) < 6 * sd * sd * np.sqrt(2.0 / N)
    zmax = ((x - mean).abs().max() / sd).item()
    # max of 2^28 |N(0,1)|: ~6.1 typical; beyond 7.5 would be p < 1e-4
    assert 5.0 < zmax < 7.5


@pytest.mark.parametrize("rbits", [32, 64])
@pytest.mark.parametrize("kind", [1, 2, 3])
def test_uniform_bits_fills(dev, kind, rbits):
    """rand(rng, UniformBitsDistribution<U>, n, r) for U of 16, 32 and 64
    bits bit-exact against the oracle (pinned to the compiled reference),
    and the reference's own determinism check
    (example/random/include/random_rng.hpp:234-305): for random sizes K,
    K scalar draws equal one bulk draw of K values, for the raw result and
    all three widths."""
    import torch
    m = _m()
    tdt = {16: torch.int16, 32: torch.int32, 64: torch.int64}
    for ubits in (16, 32, 64):
        for pre, n in ((0, 5000), (3, 4097), (1, 1), (2, 9001)):
            key = [pre + 5, 6, 7, 8][:2 if kind == 1 else 4]
            rng = _cls(kind, rbits)(key=key)
            e = orc.engine(kind, key, rbits)
            _predraw(m, rng, e, pre, rbits, dev)
            out = torch.full((n,), -1, dtype=tdt[ubits], device=dev)
            m.rand(rng, m.UniformBitsDistribution(ubits), n, out)
            want = orc.rand_uniform_bits(e, ubits, n)
            got = out.cpu().numpy().view(want.dtype)
            assert np.array_equal(got, want), (ubits, pre, n)
            assert rng.ctr() == orc.state(e)[0] and \
                rng.index == orc.state(e)[1]
    # scalar draws (engine operator()) against bulk draws, sizes from the
    # engine as in random_rng_d
    r1 = _cls(kind, rbits)(key=[9, 9, 9, 9][:2 if kind == 1 else 4])
    r2 = r1.copy()
    sizes = np.random.default_rng(kind * rbits).integers(0, 300, 12)
    mask = (1 << rbits) - 1
    for K in sizes:
        K = int(K)
        scal = [r1() & mask for _ in range(K)]
        out = _tensor(K, rbits // 8, dev)
        m.rand(r2, K, out)
        assert [int(v) for v in _u(out)[:K]] == scal
        for ubits in (16, 32, 64):
            if ubits >= rbits:
                per = ubits // rbits
                words = [r1() & mask for _ in range(K * per)]
                scal = [sum(words[i * per + j] << (rbits * j)
                            for j in range(per)) for i in range(K)]
            else:
                scal = [(r1() & mask) & ((1 << ubits) - 1) for _ in range(K)]
            out = torch.full((max(K, 1),), 0, dtype=tdt[ubits], device=dev)
            m.rand(r2, m.UniformBitsDistribution(ubits), K, out)
            a = out.cpu().numpy()[:K]
            a = a.view({16: np.uint16, 32: np.uint32, 64: np.uint64}[ubits])
            assert [int(v) for v in a] == scal, (K, ubits)


def test_wrappers_check_tensors(dev):
    """Element size, element type, contiguity and length are checked before
    a launch, and the engine does not advance."""
    import torch
    m = _m()
    rng = m.Philox4x32()
    f32 = torch.empty(8, dtype=torch.float32, device=dev)
    with pytest.raises(m.MCKLError, match="8"):
        m.rand(rng, m.NormalDistribution(0.0, 1.0, "float64"), 8, f32)
    with pytest.raises(m.MCKLError, match="integer"):
        m.rand(rng, 8, f32)
    with pytest.raises(m.MCKLError, match="8-byte|4-byte"):
        m.rand(m.Philox4x32_64(), 4, torch.empty(8, dtype=torch.int32,
                                                  device=dev))
    with pytest.raises(m.MCKLError, match="contiguous"):
        m.rand(rng, 4, torch.empty((4, 2), dtype=torch.int32,
                                   device=dev)[:, 0])
    with pytest.raises(m.MCKLError, match="needs 9"):
        m.rand(rng, m.U01OODistribution("float32"), 9, f32)
    w = torch.full((8,), 0.125, dtype=torch.float64, device=dev)
    with pytest.raises(m.MCKLError, match="expected 4"):
        m.ResampleSystematic()(8, 8, rng, w,
                               torch.empty(8, dtype=torch.int64, device=dev))
    with pytest.raises(m.MCKLError, match="expected 8"):
        m.ResampleSystematic()(8, 8, rng, w.float(),
                               torch.empty(8, dtype=torch.int32, device=dev))
    assert rng.ctr() == [0, 0] and rng.index == 4


def test_engine_errors():
    m = _m()
    e = m._lib.Engine()
    k = (ctypes.c_uint64 * 4)(1, 0, 0, 0)
    assert m.lib.mckl_hip_engine_init(ctypes.byref(e), 9, 32, k) == m._lib.EINVAL
    assert m.lib.mckl_hip_engine_init(ctypes.byref(e), 1, 16, k) == m._lib.EINVAL
    rng = m.Philox4x32()
    with pytest.raises(m.MCKLError):
        m.rand(rng, m.NormalDistribution(0.0, -1.0, "float32"), 4,
               __import__("torch").empty(4, device="cuda"))


def test_engine_text_state(dev):
    """operator<< / operator>> (random/counter.hpp:235-271) against the
    compiled reference's text after mixed call sequences, and a round trip
    that continues the stream."""
    m = _m()
    cases = json.load(open(os.path.join(GOLDEN, "engine_text.json")))
    for c in cases:
        rng = _cls(c["kind"], c["rbits"])(key=c["key"])
        for s in c["calls"]:
            if s > 0:
                out = _tensor(s, c["rbits"] // 8, dev)
                m.rand(rng, s, out)
            elif s < 0:
                rng.discard(-s)
            else:
                rng()
        assert rng.state_string() == c["text"], c
        other = _cls(c["kind"], c["rbits"])(key=[1])
        other.load_state_string(c["text"])
        assert other == rng
        assert other() == rng()
