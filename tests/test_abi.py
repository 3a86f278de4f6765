"""The C-ABI library loads and exports every entry point include/*.h declares;
host-only argument validation returns the documented status codes (no GPU
work is started by these calls)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for mt in re.finditer(r"\b(mckl_hip_[a-z0-9_]+)\s*\(", src):
            names.add(mt.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import mckl_amd
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(mckl_amd.lib, n)]
    assert not missing, missing
    assert not mckl_amd._lib.MISSING, mckl_amd._lib.MISSING


def test_engine_state_machine_host_only():
    import mckl_amd as m
    e = m._lib.Engine()
    k = (ctypes.c_uint64 * 4)(1, 0, 0, 0)
    assert m.lib.mckl_hip_engine_init(ctypes.byref(e), 1, 32, k) == 0
    assert e.index == 4 and list(e.ctr) == [0, 0, 0, 0]
    assert m.lib.mckl_hip_engine_words(ctypes.byref(e)) == 4
    # invalid kind / width
    assert m.lib.mckl_hip_engine_init(ctypes.byref(e), 7, 32, k) == m._lib.EINVAL
    assert m.lib.mckl_hip_engine_init(ctypes.byref(e), 2, 8, k) == m._lib.EINVAL
    assert b"rbits" in m.lib.mckl_hip_last_error()
    assert m.lib.mckl_hip_engine_init(None, 1, 32, k) == m._lib.EINVAL
    # Threefry4x64 with 64-bit results: 4 words per block
    assert m.lib.mckl_hip_engine_init(ctypes.byref(e), 2, 64, k) == 0
    assert m.lib.mckl_hip_engine_words(ctypes.byref(e)) == 4


@pytest.mark.parametrize("kind,rbits", [(1, 32), (2, 32), (3, 32), (1, 64)])
def test_discard_matches_oracle(kind, rbits):
    """discard() is host arithmetic (counter.hpp:178-199), including the
    strict '<' carry test of increment(ctr, n) (increment_generic.hpp:
    208-217)."""
    import mckl_amd as m
    import oracle as orc
    for start, skips in (([0, 0], [1, 5, 1000, 3]),
                         ([2 ** 64 - 1, 4], [8, 8, 8]),
                         ([2 ** 64 - 9, 0], [33, 1, 2 ** 40])):
        e = m._lib.Engine()
        k = (ctypes.c_uint64 * 4)(3, 4, 5, 6)
        m.check(m.lib.mckl_hip_engine_init(ctypes.byref(e), kind, rbits, k))
        c = (ctypes.c_uint64 * 4)(*(start + [0, 0]))
        m.check(m.lib.mckl_hip_engine_set_ctr(ctypes.byref(e), c))
        o = orc.engine(kind, [3, 4, 5, 6], rbits)
        orc.O.mo_engine_set_ctr(ctypes.byref(o), c)
        for s in skips:
            m.check(m.lib.mckl_hip_engine_discard(ctypes.byref(e), s))
            orc.O.mo_discard(ctypes.byref(o), ctypes.c_uint64(s))
            n = 4 if kind == 2 else 2
            assert list(e.ctr[:n]) == list(o.ctr[:n])
            assert e.index == o.index


def test_null_and_invalid_arguments():
    import mckl_amd as m
    e = m._lib.Engine()
    k = (ctypes.c_uint64 * 4)(1, 0, 0, 0)
    m.check(m.lib.mckl_hip_engine_init(ctypes.byref(e), 1, 32, k))
    # U01 interval out of range
    assert m.lib.mckl_hip_rand_u01_f64(ctypes.byref(e), 9, 10, None,
                                       None) == m._lib.EINVAL
    # non-positive stddev: normal_distribution.hpp:45-48
    assert m.lib.mckl_hip_rand_normal_f64(ctypes.byref(e), 10, 0.0, 0.0, None,
                                          None) == m._lib.EINVAL
    # n == 0 is a no-op, output may be NULL
    assert m.lib.mckl_hip_rand(ctypes.byref(e), 0, None, None) == 0
    assert e.index == 4


def test_resample_index_bookkeeping_host_only():
    """ResampleIndex bookkeeping and its range checks (resample.hpp:606-715)
    with empty vectors only, so no device memory is touched."""
    import mckl_amd as m
    L = m.lib
    h = ctypes.c_void_p()
    m.check(L.mckl_hip_resample_index_create(ctypes.byref(h)))
    n, r, c = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    SMAX = 2 ** 64 - 1
    for _ in range(3):
        m.check(L.mckl_hip_resample_index_push_back(h, 0, None, None))
    m.check(L.mckl_hip_resample_index_num_iter(h, ctypes.byref(n)))
    assert n.value == 3
    m.check(L.mckl_hip_resample_index_size(h, SMAX, ctypes.byref(n)))
    assert n.value == 0
    m.check(L.mckl_hip_resample_index_matrix_shape(h, 2, 1, ctypes.byref(r),
                                                   ctypes.byref(c)))
    assert (r.value, c.value) == (0, 2)
    E = m._lib.EINVAL
    assert L.mckl_hip_resample_index_matrix_shape(
        h, 3, 0, ctypes.byref(r), ctypes.byref(c)) == E
    assert b"out of range" in L.mckl_hip_last_error()
    assert L.mckl_hip_resample_index_index(h, 0, None, 1, 2, None, None) == E
    assert L.mckl_hip_resample_index_read_matrix(h, 5, 2, 0, None, None) == E
    # n == 0 ids and N == 0 rows are no-ops
    assert L.mckl_hip_resample_index_index(h, 0, None, 2, 0, None, None) == 0
    assert L.mckl_hip_resample_index_read_matrix(h, 101, 2, 0, None,
                                                 None) == 0
    m.check(L.mckl_hip_resample_index_reset(h))
    m.check(L.mckl_hip_resample_index_num_iter(h, ctypes.byref(n)))
    assert n.value == 0
    m.check(L.mckl_hip_resample_index_size(h, 1, ctypes.byref(n)))
    assert n.value == 0
    m.check(L.mckl_hip_resample_index_clear(h))
    m.check(L.mckl_hip_resample_index_push_back(h, 0, None, None))
    m.check(L.mckl_hip_resample_index_destroy(h))


def test_resample_rejects_null_weights():
    """N > 1 with M > 0 reads the weights: a NULL weight pointer is
    MCKL_HIP_EINVAL before any launch (N == 1 and M == 0 never read them)."""
    import mckl_amd as m
    e = m._lib.Engine()
    k = (ctypes.c_uint64 * 4)(1, 0, 0, 0)
    m.check(m.lib.mckl_hip_engine_init(ctypes.byref(e), 1, 32, k))
    host = (ctypes.c_uint32 * 8)()  # never reached by a kernel
    for fn in (m.lib.mckl_hip_resample, m.lib.mckl_hip_resample_noalloc):
        for scheme in range(6):
            rc = fn(scheme, 5, 5, ctypes.byref(e), None, host, None, None)
            assert rc == m._lib.EINVAL
            assert b"NULL weights" in m.lib.mckl_hip_last_error()
    assert list(e.ctr[:2]) == [0, 0]


def test_python_wrappers_reject_host_tensors():
    """Every wrapper checks its tensors before a launch: host memory,
    non-tensors and short tensors raise MCKLError."""
    import torch
    import mckl_amd as m
    rng = m.Philox4x32()
    cpu = torch.zeros(16, dtype=torch.int32)
    with pytest.raises(m.MCKLError, match="not in GPU memory"):
        m.rand(rng, 16, cpu)
    with pytest.raises(m.MCKLError, match="expected a torch tensor"):
        m.rand(rng, 16, [0] * 16)
    with pytest.raises(m.MCKLError, match="not in GPU memory"):
        m.rand(rng, m.NormalDistribution(0.0, 1.0, "float32"), 4,
               torch.zeros(4, dtype=torch.float32))
    with pytest.raises(m.MCKLError, match="not in GPU memory"):
        m.rand(rng, m.U01OODistribution("float64"), 4,
               torch.zeros(4, dtype=torch.float64))
    with pytest.raises(m.MCKLError, match="not in GPU memory"):
        m.ResampleSystematic()(4, 4, rng, torch.full((4,), 0.25,
                                                     dtype=torch.float64),
                               torch.zeros(4, dtype=torch.int32))
    with pytest.raises(m.MCKLError, match="not in GPU memory"):
        m.resample_trans_rep_index(4, 4, torch.ones(4, dtype=torch.int32),
                                   torch.zeros(4, dtype=torch.int32))
    assert rng.ctr() == [0, 0] and rng.index == 4
