"""GPU parity of the SeedGenerator keys and the RNG-set fills against the
oracle (itself pinned to the reference's random_seed_*.data KAT files)."""
import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

KINDS = [orc.PHILOX4X32, orc.THREEFRY4X64, orc.ARS]


def _m():
    import mckl_amd as m
    return m


def _cls(m, kind):
    return {orc.PHILOX4X32: m.Philox4x32, orc.THREEFRY4X64: m.Threefry4x64,
            orc.ARS: m.ARS}[kind]


def _u64(t):
    return t.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("randomize", [True, False])
def test_seed_keys(dev, kind, randomize):
    m = _m()
    sg = m.SeedGenerator(_cls(m, kind), randomize)
    got = np.concatenate([_u64(sg.keys(3000)), _u64(sg.keys(17))])
    want = orc.seed_keys(kind, 1, 3017, randomize=randomize)
    assert np.array_equal(got, want)
    sg = m.SeedGenerator(_cls(m, kind), randomize)
    sg.set(123456789012)
    sg.partition(7, 5)
    got = _u64(sg.keys(1000))
    want = orc.seed_keys(kind, 123456789012, 1000, np_=7, rank=5,
                         randomize=randomize)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind", KINDS)
def test_rngset_fills(dev, kind):
    import torch
    m = _m()
    cls = _cls(m, kind)
    n = 300
    rs = m.RNGSetVector(cls, n)
    keys = orc.seed_keys(kind, 1, n)
    for per in (1, 5, 37):
        out = torch.zeros(n * per, dtype=torch.int32, device=dev)
        rs.rand(per, out)
        got = out.cpu().numpy().view(np.uint32).reshape(n, per)
        for e in (0, 1, 150, n - 1):
            k = [int(v) for v in keys[e]][:cls.KEY_WORDS]
            want = orc.rand_u32(orc.engine(kind, k), per)
            assert np.array_equal(got[e], want), (per, e)
        for dt in (torch.float64, torch.float32):
            f = torch.zeros(n * per, dtype=dt, device=dev)
            rs.u01(per, f)
            fg = f.cpu().numpy().reshape(n, per)
            npdt = np.float64 if dt == torch.float64 else np.float32
            for e in (0, 77, n - 1):
                k = [int(v) for v in keys[e]][:cls.KEY_WORDS]
                want = orc.rand_u01(orc.engine(kind, k), orc.OO, per, npdt)
                assert np.array_equal(fg[e], want), (per, e, dt)
    # a member engine continues its stream on the host side
    e0 = rs.engine(3)
    k = [int(v) for v in keys[3]][:cls.KEY_WORDS]
    assert e0.key() == k


def test_seed_errors(dev):
    m = _m()
    with pytest.raises(m.MCKLError):
        m.SeedGenerator().partition(2, 2)
