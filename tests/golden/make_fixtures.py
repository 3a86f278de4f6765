"""Write the committed fixtures of tests/golden from the compiled reference.

Runs in the container that mounts the reference: it loads
oracle/_ref/libmckl_ref.so (the reference's own headers behind
oracle/ref_driver.cpp, built by `make -C oracle ref`) and records input and
output vectors only.  Re-run with `python tests/golden/make_fixtures.py`.

Fixtures:
  config1.json         Philox4x32 key {0x89ABCDEF, 0x01234567}, counter 0:
                       SHA-256 of the first 2^24 u32 words and of the same
                       stream converted by U01OO<double>; head/tail windows.
  normal_streams.npz   {Threefry4x64, ARS} x {Normal<float>(0,1),
                       Normal<double>(2.5,0.3)}, seeds 1, 2, 64, 2^22 values:
                       three windows each plus the final engine state.
  resample_cases.npz   W, rep, idx from ResampleAlgorithm + rep->index for the
                       six schemes (ResampleScheme numbering) at several
                       (N, M); u is the scheme's uniform sequence over M.
  weight_cases.npz     Weight::set_log / add_log inputs and outputs.
  observations.npz     simulated observations of the SMC configurations (the
                       recipes of mckl_amd/workloads.py on reference draws).
  engine_text.json     CounterEngine's textual state (operator<<) after call
                       sequences (bulk n > 0, scalar 0, discard -n), all six
                       engines.

`python tests/golden/make_fixtures.py NAME ...` writes only the named ones.
"""
import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import oracle as orc  # noqa: E402

KEY1 = [0x89ABCDEF, 0x01234567]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def config1():
    n = 1 << 24
    u32, ctr, idx = orc.ref_engine_ops(orc.PHILOX4X32, 32, KEY1, [n])
    f64, ctr2, idx2 = orc.ref_u01(orc.PHILOX4X32, KEY1, orc.OO, np.float64, n)
    d = {
        "key": KEY1, "n": n,
        "sha256_u32": sha(u32), "sha256_u01oo_f64": sha(f64),
        "head_u32": [int(v) for v in u32[:64]],
        "tail_u32": [int(v) for v in u32[-64:]],
        "head_u01oo_f64_hex": [float(v).hex() for v in f64[:16]],
        "state_after_u32": {"ctr": ctr, "index": idx},
        "state_after_u01": {"ctr": ctr2, "index": idx2},
    }
    with open(os.path.join(HERE, "config1.json"), "w") as f:
        json.dump(d, f, indent=1)


def normal_streams():
    n = 1 << 22
    out = {}
    for ename, kind in (("threefry", orc.THREEFRY4X64), ("ars", orc.ARS)):
        for dname, dt, mean, sd in (("f32", np.float32, 0.0, 1.0),
                                    ("f64", np.float64, 2.5, 0.3)):
            for s in (1, 2, 64):
                r, ctr, idx = orc.ref_normal(kind, [s], dt, n, mean, sd)
                p = "%s_%s_%d" % (ename, dname, s)
                out[p + "_head"] = r[:4096]
                out[p + "_mid"] = r[n // 2 - 2048:n // 2 + 2048]
                out[p + "_tail"] = r[-4096:]
                out[p + "_state"] = np.array(ctr + [idx], np.uint64)
    np.savez_compressed(os.path.join(HERE, "normal_streams.npz"), **out)


def resample_cases():
    rng = np.random.default_rng(20240601)
    out = {}
    k = 0
    for scheme in (orc.MULTINOMIAL, orc.STRATIFIED, orc.SYSTEMATIC,
                   orc.RESIDUAL, orc.RESIDUAL_STRATIFIED,
                   orc.RESIDUAL_SYSTEMATIC):
        for N, M in ((1, 1), (2, 2), (3, 3), (1000, 1000), (4097, 4097),
                     (300, 700), (700, 300), (65536, 65536)):
            w = rng.random(N) ** 4
            w /= w.sum()
            key = [1000 + k, 0]
            rep, idx = orc.ref_resample(scheme, orc.PHILOX4X32, key, w, M)
            u = orc.ref_u01_sequence(scheme, orc.PHILOX4X32, key, M)
            p = "c%d" % k
            out[p + "_meta"] = np.array([scheme, N, M, key[0]], np.int64)
            out[p + "_w"] = w
            out[p + "_u"] = u
            out[p + "_rep"] = rep
            out[p + "_idx"] = idx
            k += 1
    np.savez_compressed(os.path.join(HERE, "resample_cases.npz"), **out)


def weight_cases():
    rng = np.random.default_rng(7)
    out = {}
    for k, n in enumerate((1, 2, 1023, 1024, 1025, 5000, 100000)):
        v = rng.normal(size=n) * 4.0
        w, ess = orc.ref_weight(0, v)
        v2 = rng.normal(size=n) * 2.0
        w2, ess2 = orc.ref_weight(1, v2, w, ess)
        out["w%d_v" % k] = v
        out["w%d_W" % k] = w
        out["w%d_ess" % k] = np.array([ess])
        out["w%d_v2" % k] = v2
        out["w%d_W2" % k] = w2
        out["w%d_ess2" % k] = np.array([ess2])
    np.savez_compressed(os.path.join(HERE, "weight_cases.npz"), **out)


def observations():
    from mckl_amd_workloads_ref import lg_from_z, sv_from_z, sv_params_from_u
    out = {}
    z, _, _ = orc.ref_normal(orc.PHILOX4X32, [7, 0], np.float64, 201, 0.0, 1.0)
    out["lg_y"] = lg_from_z(z, 100)
    z, _, _ = orc.ref_normal(orc.PHILOX4X32, [11, 0], np.float64, 1001, 0.0, 1.0)
    out["sv_y"] = sv_from_z(z, 500, 0.95, 0.2, 0.7)
    th, nf, ys = [], [], []
    for f in range(64):
        u, _, _ = orc.ref_u01(orc.PHILOX4X32, [f, 0xB47C4], orc.OO,
                              np.float64, 4)
        z, _, _ = orc.ref_normal(orc.PHILOX4X32, [f, 0xB47C4], np.float64,
                                 101, 0.0, 1.0, pre=4)
        phi, s, beta, n = sv_params_from_u(u)
        th.append((phi, s, beta))
        nf.append(n)
        ys.append(sv_from_z(z, 50, phi, s, beta))
    out["batch64_theta"] = np.array(th)
    out["batch64_N"] = np.array(nf, np.uint32)
    out["batch64_y"] = np.array(ys)
    np.savez_compressed(os.path.join(HERE, "observations.npz"), **out)


ENGINE_TEXT_SEQS = [
    [5, 0, -3], [1], [4097], [0], [3, -100, 2], [-1], [4], [8, 0, 0, 0],
    [2, -(1 << 40), 1], [1000, -7, 0, 11],
]


def engine_text():
    import ctypes
    cases = []
    for kind, rbits, key in ((orc.PHILOX4X32, 32, [7, 9]),
                             (orc.PHILOX4X32, 64, [0xDEADBEEF, 3]),
                             (orc.THREEFRY4X64, 32, [1, 2, 3, 4]),
                             (orc.THREEFRY4X64, 64, [2 ** 63 + 5, 0, 7, 1]),
                             (orc.ARS, 32, [5, 6, 7, 8]),
                             (orc.ARS, 64, [0xFFFFFFFF, 1, 2, 3])):
        for seq in ENGINE_TEXT_SEQS:
            k = (ctypes.c_uint64 * 4)(*(key + [0] * (4 - len(key))))
            sz = (ctypes.c_longlong * len(seq))(*seq)
            buf = ctypes.create_string_buffer(4096)
            n = orc.REF.ref_engine_text(kind, rbits, k, len(seq), sz, buf,
                                        4096)
            assert n > 0
            cases.append({"kind": kind, "rbits": rbits, "key": key,
                          "calls": seq, "text": buf.value.decode()})
    with open(os.path.join(HERE, "engine_text.json"), "w") as f:
        json.dump(cases, f, indent=1)


if __name__ == "__main__":
    if len(sys.argv) > 1:
        if orc.REF is None:
            sys.exit("oracle/_ref/libmckl_ref.so is not built")
        for name in sys.argv[1:]:
            globals()[name]()
        sys.exit(0)
    if orc.REF is None:
        sys.exit("oracle/_ref/libmckl_ref.so is not built (make -C oracle ref)")
    # the workload recipes, imported without importing the GPU package
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mckl_amd_workloads_ref", os.path.join(ROOT, "mckl_amd", "workloads.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    sys.modules["mckl_amd_workloads_ref"] = mod
    config1()
    normal_streams()
    resample_cases()
    weight_cases()
    observations()
    print("fixtures written to", HERE)
