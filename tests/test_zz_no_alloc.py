"""mckl_hip_resample_noalloc (workspace in a device array of the library)
against mckl_hip_resample (stream-ordered allocation), through the C ABI:
replication counts, parent indices and the engine advance bit for bit, two
calls in a row with the second one smaller, for all six schemes."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NMAX = 1 << 20


def _engine(m, key):
    e = m._lib.Engine()
    k = (ctypes.c_uint64 * 4)(key, 0, 0, 0)
    m.check(m.lib.mckl_hip_engine_init(ctypes.byref(e), m._lib.PHILOX4X32, 32,
                                       k))
    return e


def _weights(torch, dev, N, seed):
    rng = np.random.default_rng(seed)
    w = rng.random(N) ** 4
    w /= w.sum()
    return torch.from_numpy(w).to(dev)


def _run(m, torch, dev, fn, scheme, N, M, w, key):
    e = _engine(m, key)
    rep = torch.full((N,), -1, dtype=torch.int32, device=dev)
    idx = torch.full((M,), -1, dtype=torch.int32, device=dev)
    m.check(fn(scheme, N, M, ctypes.byref(e), ctypes.c_void_p(w.data_ptr()),
               ctypes.c_void_p(rep.data_ptr()),
               ctypes.c_void_p(idx.data_ptr()), None))
    torch.cuda.synchronize()
    return (rep.cpu().numpy(), idx.cpu().numpy(),
            (list(e.ctr[:2]), e.index))


@pytest.mark.parametrize("scheme", [0, 1, 2, 3, 4, 5])
def test_noalloc_matches_allocating_form(dev, scheme):
    import torch
    import mckl_amd as m
    for N, M, key in ((NMAX, NMAX, 11), (5000, 4097, 12)):
        print("resample_noalloc scheme %d N %d M %d" % (scheme, N, M),
              flush=True)
        w = _weights(torch, dev, N, N * 7 + scheme)
        want = _run(m, torch, dev, m.lib.mckl_hip_resample, scheme, N, M, w,
                    key)
        got = _run(m, torch, dev, m.lib.mckl_hip_resample_noalloc, scheme, N,
                   M, w, key)
        assert np.array_equal(got[0], want[0]), (scheme, N, M)
        assert np.array_equal(got[1], want[1]), (scheme, N, M)
        assert got[2] == want[2]
        assert int(got[0].astype(np.int64).sum()) == M


def test_noalloc_rejects_sizes_beyond_the_workspace(dev):
    import torch
    import mckl_amd as m
    e = _engine(m, 1)
    w = torch.full((4,), 0.25, dtype=torch.float64, device=dev)
    rep = torch.zeros(4, dtype=torch.int32, device=dev)
    for N, M in ((NMAX + 1, 4), (4, NMAX + 1)):
        rc = m.lib.mckl_hip_resample_noalloc(
            2, N, M, ctypes.byref(e), ctypes.c_void_p(w.data_ptr()),
            ctypes.c_void_p(rep.data_ptr()), None, None)
        assert rc == m._lib.EINVAL
        assert b"workspace limit" in m.lib.mckl_hip_last_error()
    assert list(e.ctr[:2]) == [0, 0]
