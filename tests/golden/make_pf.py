"""Writes pf_reference.npz: runs of the reference's own particle-filter
example (example/algorithm/include/algorithm_pf.hpp:38-190, BackendSEQ,
Stratified resampling at threshold 0.5, no estimators) through
oracle/_ref/libmckl_ref_pf.so, on the example's observations
(algorithm_pf.data, copied beside this script from
example/algorithm/algorithm_pf.data).

Per case (N, SeedGenerator internal seed): the ESS and size histories, the
final state (column-major p, q, s, t) and normalised weights -- all N rows
up to N = 10000, every 100th row at the example's own N = 100000 -- and the
key words of the RNG set's engine 0 and of Particle::rng().

Run from the repository root, where the reference is mounted:
python tests/golden/make_pf.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as orc  # noqa: E402

LIB = os.path.join(orc.HERE, "_ref", "libmckl_ref_pf.so")
CASES = ((1000, 1), (10000, 7), (100000, 1))


def run(lib, N, seed):
    T = ctypes.c_size_t()
    ess = np.zeros(1000)
    size = np.zeros(1000, np.uint64)
    state = np.zeros(4 * N)
    w = np.zeros(N)
    keys = np.zeros(4, np.uint64)
    rc = lib.ref_pf(ctypes.c_size_t(N), ctypes.c_uint64(seed),
                    HERE.encode(), ctypes.c_size_t(1000), ctypes.byref(T),
                    orc.P(ess), orc.P(size), orc.P(state), orc.P(w),
                    orc.P(keys))
    assert rc == 0, "ref_pf failed"
    t = T.value
    return ess[:t], size[:t], state.reshape(4, N), w, keys


def main():
    lib = ctypes.CDLL(LIB, mode=os.RTLD_LAZY)
    out = {}
    for N, seed in CASES:
        ess, size, state, w, keys = run(lib, N, seed)
        rows = np.arange(0, N, 100 if N > 10000 else 1)
        tag = "N%d_s%d" % (N, seed)
        out[tag + "_ess"] = ess
        out[tag + "_size"] = size
        out[tag + "_rows"] = rows
        out[tag + "_state"] = state[:, rows]
        out[tag + "_weight"] = w[rows]
        out[tag + "_keys"] = keys
        print(tag, "T", len(ess), "resampled", int((ess < 0.5 * N).sum()))
    np.savez_compressed(os.path.join(HERE, "pf_reference.npz"), **out)


if __name__ == "__main__":
    main()
