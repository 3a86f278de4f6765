"""Benchmark of the MI355X Monte Carlo hot path (one JSON line on stdout).

Headline (value): Philox4x32-10 bulk-fill throughput on the raw-fill
configuration -- key 0x0123456789ABCDEF (key_type {0x89ABCDEF, 0x01234567}),
counter 0, 2^24 raw u32 words via rand(rng, n, r) and the same stream
converted to 2^24 U01OO doubles; one step = both fills, value = bytes
written per second.  Outputs stay in HBM; nothing is copied to the host
inside the timed region.

Also reported (in "secondary"): the 2^28-word saturation fill, the
distribution fills (Threefry4x64 / ARS x Normal<float>(0,1) /
Normal<double>(2.5,0.3), 64 streams x 2^22), the plain SMC on the
linear-Gaussian configuration (N = 2^20, T = 100), and the batched SMC on
4096 stochastic-volatility filters (T = 50) against one plain launch
sequence per filter.

roofline: the dominant kernel (the U01OO<double> fill) timed with HIP
events on its own stream inside the timed region; algorithmic bytes =
8 bytes per value written.  cpu_baseline: the reference compiled from its
own headers (oracle/_ref) or else the oracle port, on this host, one thread.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KEY1 = [0x89ABCDEF, 0x01234567]


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_config1(m, torch, steps, warmup):
    n = 1 << 24
    u32 = torch.empty(n, dtype=torch.int32, device="cuda")
    f64 = torch.empty(n, dtype=torch.float64, device="cuda")
    dist = m.U01OODistribution("float64")
    k_raw = []
    k_u01 = []

    def step(record):
        rng = m.Philox4x32(key=KEY1)
        if record:
            a, b = _events()
            a.record()
        m.rand(rng, n, u32)
        if record:
            b.record()
            k_raw.append((a, b))
        rng = m.Philox4x32(key=KEY1)
        if record:
            a, b = _events()
            a.record()
        m.rand(rng, dist, n, f64)
        if record:
            b.record()
            k_u01.append((a, b))

    for _ in range(warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    raw_ms = sum(a.elapsed_time(b) for a, b in k_raw) / steps
    u01_ms = sum(a.elapsed_time(b) for a, b in k_u01) / steps
    return {
        "n": n, "wall_s": wall, "bytes_per_step": 12 * n,
        "raw_ms": raw_ms, "u01_ms": u01_ms,
    }


def bench_saturation(m, torch, reps=10):
    n = 1 << 28
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    res = {}
    for name, cls, key in (("philox4x32", m.Philox4x32, KEY1),
                           ("threefry4x64", m.Threefry4x64, [1, 0, 0, 0]),
                           ("ars", m.ARS, [1, 0, 0, 0])):
        rng = cls(key=key)
        m.rand(rng, n, out)
        a, b = _events()
        a.record()
        for _ in range(reps):
            m.rand(rng, n, out)
        b.record()
        torch.cuda.synchronize()
        t = a.elapsed_time(b) / reps / 1e3
        res[name + "_u32_2^28_GBps"] = round(4 * n / t / 1e9, 1)
    del out
    return res


def bench_normals(m, torch, reps=3):
    n = 1 << 22
    res = {}
    for ename, cls in (("threefry4x64", m.Threefry4x64), ("ars", m.ARS)):
        for dname, dt, mean, sd in (("f32", torch.float32, 0.0, 1.0),
                                    ("f64", torch.float64, 2.5, 0.3)):
            out = torch.empty(64 * n, dtype=dt, device="cuda")
            engines = [cls(key=[s, 0, 0, 0]) for s in range(1, 65)]
            m.streams_normal(engines, n, mean, sd, out)
            a, b = _events()
            a.record()
            for _ in range(reps):
                engines = [cls(key=[s, 0, 0, 0]) for s in range(1, 65)]
                m.streams_normal(engines, n, mean, sd, out)
            b.record()
            torch.cuda.synchronize()
            t = a.elapsed_time(b) / reps / 1e3
            res["normal_%s_%s_2^28" % (ename, dname)] = {
                "Gsamples_per_s": round(64 * n / t / 1e9, 2),
                "GBps": round(out.element_size() * 64 * n / t / 1e9, 1),
                "ms": round(t * 1e3, 3)}
            del out
    return res


def bench_smc(m, torch):
    from mckl_amd import workloads as wl
    import numpy as np
    res = {}
    # plain SMC, linear-Gaussian, N = 2^20, T = 100, systematic
    y = wl.lg_observations(100, 7)
    N, T = 1 << 20, 100
    s = m.SMCSampler(m.MODEL_LG, (0.9, 1.0, 0.5), y, N, 1, 0.5, m.SYSTEMATIC)
    s.iterate()
    torch.cuda.synchronize()
    a, b = _events()
    reps = 5
    a.record()
    for _ in range(reps):
        s.iterate()
    b.record()
    torch.cuda.synchronize()
    t = a.elapsed_time(b) / reps / 1e3
    ll, ess, rs = s.results()
    res["smc_plain_lg_N2^20_T100"] = {
        "particle_steps_per_s": N * T / t, "ms": round(t * 1e3, 3),
        "loglik": ll, "resampling_steps": int(rs.sum())}

    # batched SV filters
    theta, Nf, yb, seeds = wl.sv_batch(4096, 50)
    T = 50
    ps = float(Nf.astype(np.float64).sum()) * T
    bt = m.SMCBatch(m.MODEL_SV, Nf, theta, yb, seeds, 0.5, m.STRATIFIED)
    best = None
    tune_log = []
    for block in (256, 512):
        for wpe in (0, 2):
            bt.set_tune(block, wpe)
            bt.run(use_graph=True)
            torch.cuda.synchronize()
            a, b = _events()
            a.record()
            bt.run(use_graph=True)
            b.record()
            torch.cuda.synchronize()
            tt = a.elapsed_time(b) / 1e3
            tune_log.append({"block": block, "waves_per_eu": wpe,
                             "ms": round(tt * 1e3, 3)})
            if best is None or tt < best[0]:
                best = (tt, block, wpe)
    bt.set_tune(best[1], best[2])
    bt.run(use_graph=True)
    torch.cuda.synchronize()
    a, b = _events()
    a.record()
    reps = 3
    for _ in range(reps):
        bt.run(use_graph=True)
    b.record()
    torch.cuda.synchronize()
    tb = a.elapsed_time(b) / reps / 1e3
    lb = bt.log_likelihood()

    # one plain launch sequence per filter, same GPU
    samplers = [m.SMCSampler(m.MODEL_SV, theta[f], yb[f], int(Nf[f]),
                             int(seeds[f]), 0.5, m.STRATIFIED)
                for f in range(len(Nf))]
    torch.cuda.synchronize()
    a, b = _events()
    a.record()
    for sm in samplers:
        sm.iterate()
    b.record()
    torch.cuda.synchronize()
    tp = a.elapsed_time(b) / 1e3
    lp = np.array([sm.log_likelihood() for sm in samplers[:64]])
    res["smc_batched_sv_4096_T50"] = {
        "particle_steps": ps,
        "batched_particle_steps_per_s": ps / tb,
        "per_filter_particle_steps_per_s": ps / tp,
        "speedup_vs_per_filter": tp / tb,
        "batched_ms": round(tb * 1e3, 3), "per_filter_ms": round(tp * 1e3, 1),
        "tune": {"block": best[1], "waves_per_eu": best[2], "sweep": tune_log},
        "bit_identical_first64": bool(np.array_equal(lb[:64], lp)),
    }
    return res


def cpu_baseline(steps=10):
    """The reference's own bulk fills (compiled from its headers), one
    thread; else the oracle port.  Sample: the full config-1 step."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    n = 1 << 24
    u32 = np.empty(n, np.uint32)
    f64 = np.empty(n, np.float64)
    if orc.REF is not None:
        key = (ctypes.c_uint64 * 4)(*(KEY1 + [0, 0]))
        t = orc.REF.ref_time_philox_u01(key, ctypes.c_size_t(n), steps,
                                        orc.P(u32), orc.P(f64))
        kind = "reference"
    else:
        ts = []
        for _ in range(steps):
            t0 = time.perf_counter()
            e = orc.engine(orc.PHILOX4X32, KEY1)
            orc.O.mo_rand_u32(ctypes.byref(e), ctypes.c_size_t(n), orc.P(u32))
            e = orc.engine(orc.PHILOX4X32, KEY1)
            orc.O.mo_rand_u01_f64(ctypes.byref(e), 3, ctypes.c_size_t(n),
                                  orc.P(f64))
            ts.append(time.perf_counter() - t0)
        t = sorted(ts)[len(ts) // 2]
        kind = "port"
    return {"value": round(12 * n / t / 1e9, 3), "unit": "GB/s", "cores": 1,
            "kind": kind,
            "sample": "full step: 2^24 u32 + 2^24 U01OO doubles, Philox4x32, "
                      "median of %d runs, single thread" % steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-secondary", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the fills and samplers shard by "
              "counter range / filter with no communication, so no multi-GPU "
              "run is made; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import mckl_amd as m

    c1 = bench_config1(m, torch, args.steps, args.warmup)
    value = c1["bytes_per_step"] * args.steps / c1["wall_s"] / 1e9
    u01_gbs = 8 * c1["n"] / (c1["u01_ms"] / 1e3) / 1e9
    raw_gbs = 4 * c1["n"] / (c1["raw_ms"] / 1e3) / 1e9
    secondary = {"raw_u32_fill_GBps": round(raw_gbs, 1),
                 "raw_u32_fill_ms": round(c1["raw_ms"], 4),
                 "u01oo_f64_fill_GBps": round(u01_gbs, 1),
                 "u01oo_f64_fill_ms": round(c1["u01_ms"], 4)}
    if not args.no_secondary:
        secondary.update(bench_saturation(m, torch))
        secondary.update(bench_normals(m, torch))
        secondary.update(bench_smc(m, torch))
    out = {
        "metric": "Philox GB/s vs HBM peak; SMC particle-steps/s plain vs "
                  "batched+adjoint",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(c1["wall_s"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32+f64",
        "data": "synthetic (counter-based streams, fixed keys)",
        "config": {
            "workload": "Philox4x32-10 raw fill 2^24 u32 + U01OO<double> "
                        "2^24 from the same stream (key 0x0123456789ABCDEF, "
                        "counter 0)",
            "n_words": c1["n"], "bytes_per_step": c1["bytes_per_step"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "fill_kernel<PhiloxGen, ConvU01F64>",
            "achieved": round(u01_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(u01_gbs / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "cpu_baseline": cpu_baseline(),
        "secondary": secondary,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
