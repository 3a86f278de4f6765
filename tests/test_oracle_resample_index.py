"""The ResampleIndex restatement in the oracle, pinned three ways: a hand
traced genealogy, the reference example's own check (trace-back equals the
direct copy of the history, example/algorithm/include/
algorithm_resample_index.hpp:37-135) with multinomial resampling between
generations of fixed and random sizes, and op-by-op agreement with the
compiled reference (oracle/_ref) over push/reset/clear scripts."""
import numpy as np
import pytest

import oracle as orc

ROW, COL = 101, 102


def genealogy(sizes, seed):
    """Index vectors as the reference example draws them
    (algorithm_resample.hpp:46-120): index[d] maps the sizes[d+1] particles
    of generation d+1 to parents among the sizes[d] of generation d."""
    rng = np.random.default_rng(seed)
    e = orc.engine(orc.PHILOX4X32, [seed, 0])
    index = []
    for d in range(len(sizes) - 1):
        w = rng.random(sizes[d])
        w /= w.sum()
        _, idx = orc.resample(orc.MULTINOMIAL, e, w, sizes[d + 1])
        index.append(np.asarray(idx, np.uint32))
    return index


def direct_copy(sizes, value, index):
    """algorithm_resample_index_rands: grow the state one column per
    generation and copy whole rows through each index vector."""
    state = np.zeros((sizes[0], 0), np.int64)
    for d, idx in enumerate(index):
        state = np.concatenate(
            [state, np.asarray(value[d], np.int64)[:, None]], axis=1)
        state = state[idx]
    return state


def test_hand_traced():
    ri = orc.ResampleIndex()
    ri.push_back(4, [0, 0, 1, 3])   # gen 1 <- gen 0
    ri.push_back(4, [3, 2, 2, 0])   # gen 2 <- gen 1
    ri.push_back(4, [1, 1, 0, 2])   # gen 3 <- gen 2
    # particle 0 of gen 3: parent 1 in gen 2, 2 in gen 1, 1 in gen 0;
    # particle 2: 0, 3, 3
    m = ri.index_matrix(ROW).reshape(4, 3)
    assert m.tolist() == [[1, 2, 1], [1, 2, 1], [3, 3, 0], [1, 2, 2]]
    assert ri.index_matrix(COL).reshape(3, 4).T.tolist() == m.tolist()
    assert ri.index([0, 3]).tolist() == [1, 1]
    assert ri.index([0, 3], 2, 1).tolist() == [2, 2]
    # partial window iter..iter_back = 1..2
    assert ri.index_matrix(ROW, 2, 1).reshape(4, 2).tolist() == \
        [[2, 1], [2, 1], [3, 0], [2, 2]]
    assert ri.index([7]).tolist() == [orc.ResampleIndex.BAD]
    with pytest.raises(ValueError):
        ri.index([0], 3, 0)
    with pytest.raises(ValueError):
        ri.index([0], 1, 2)


@pytest.mark.parametrize("fixed", [True, False])
@pytest.mark.parametrize("layout", [ROW, COL])
def test_trace_equals_direct_copy(fixed, layout):
    N, dim = 60, 45
    rng = np.random.default_rng(3)
    sizes = [N] * (dim + 1) if fixed else \
        [int(v) for v in rng.integers(N // 2, 2 * N + 1, dim)] + [N]
    value = [rng.integers(0, 2 ** 31, s) for s in sizes[:-1]]
    index = genealogy(sizes, 11 + fixed)
    ri = orc.ResampleIndex()
    for idx in index:
        ri.push_back(len(idx), idx)
    m = ri.index_matrix(layout)
    m = m.reshape(N, dim) if layout == ROW else m.reshape(dim, N).T
    traced = np.stack([np.asarray(value[j])[m[:, j]] for j in range(dim)], 1)
    assert np.array_equal(traced, direct_copy(sizes, value, index))


def _queries(n_iter, rng, sizes):
    q = []
    for _ in range(5):
        ib = int(rng.integers(0, n_iter))
        it = int(rng.integers(0, ib + 1))
        q.append(("index", rng.integers(0, sizes[-1], 7).astype(np.uint32),
                  ib, it))
        q.append(("matrix", ROW if rng.random() < 0.5 else COL, ib, it))
    return q


def _run_oracle(script):
    ri = orc.ResampleIndex()
    out = []
    for op in script:
        if op[0] == "push":
            ri.push_back(op[1], op[2])
        elif op[0] == "reset":
            ri.reset()
        elif op[0] == "clear":
            ri.clear()
        elif op[0] == "num_iter":
            out.append(ri.num_iter)
        elif op[0] == "size":
            out.append(ri.size(op[1]))
        elif op[0] == "index":
            out.append(ri.index(op[1], op[2], op[3]))
        elif op[0] == "matrix":
            out.append(ri.index_matrix(op[1], op[2], op[3]))
    return out


def build_script(seed):
    """pushes with and without data, then a reset and fewer pushes of other
    sizes (so the last vector held is stale and index() starts there), then
    a clear and a reset with fresh pushes; queries in between.  Only queries
    the reference's runtime_asserts accept are issued, all traced entries
    index their vectors: sizes are at least 8 and entries below 8 (identity
    vectors have 8 entries)."""
    rng = np.random.default_rng(seed)
    ops, sizes = [], []

    def push(k):
        for _ in range(k):
            if rng.random() < 0.25:
                ops.append(("push", 8, None))
                sizes.append(8)
            else:
                N = int(rng.integers(8, 40))
                ops.append(("push", N, rng.integers(0, 8, N).astype(np.uint32)))
                sizes.append(N)

    push(7)
    ops += [("num_iter",), ("size", None), ("size", 3), ("size", 50)]
    ops += _queries(7, rng, [8])
    ops.append(("reset",))
    sizes.clear()
    push(4)
    ops += [("num_iter",), ("size", None), ("size", 5)]
    ops += _queries(4, rng, [8])
    ops += [("clear",), ("reset",)]
    sizes.clear()
    push(5)
    ops += [("num_iter",), ("size", None)]
    ops += _queries(5, rng, [8])
    return ops


@pytest.mark.skipif(orc.REF is None, reason="reference build absent")
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_matches_compiled_reference(seed):
    script = build_script(seed)
    want = orc.ref_resample_index(script)
    got = _run_oracle(script)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert np.array_equal(np.asarray(g, np.uint64),
                              np.asarray(w, np.uint64))


def test_identity_and_stale_vector():
    ri = orc.ResampleIndex()
    ri.push_back(5)
    ri.push_back(5, [4, 3, 2, 1, 0])
    ri.push_back(5)
    ri.reset()
    ri.push_back(5, [1, 1, 1, 1, 1])
    # index() starts from the last vector held (identity, iteration 2)
    assert ri.num_iter == 1 and ri.size() == 5
    assert ri.index([3], 0, 0).tolist() == [3]
    assert ri.index_matrix(ROW, 0, 0).tolist() == [1, 1, 1, 1, 1]
    ri.clear()
    with pytest.raises(IndexError):
        ri.index([0], 0, 0)
