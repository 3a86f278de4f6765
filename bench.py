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
linear-Gaussian configuration (N = 2^20, T = 100), the same filter
written as evaluation objects of the generic SMCSampler<T>, the batched
SMC on
4096 stochastic-volatility filters (T = 50) against one plain launch
sequence per filter, the SMC log-likelihood gradient (256 points x
N = 2^16 x T = 500, with its finite-difference and Kalman checks) and the
ResampleIndex trace-back (N = 2^20 x 256 iterations).
`--only LEG` runs one of these legs alone (for profiling).

roofline: the dominant kernel (the U01OO<double> fill) timed with HIP
events on its own stream inside the timed region; algorithmic bytes =
8 bytes per value written.  cpu_baseline: the reference compiled from its
own headers (oracle/_ref) or else the oracle port, on this host, one thread.
"""
import argparse
import ctypes
import json
import math
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


def _warm(torch, fn, ms=50.0, cap=2000):
    """Calls fn until ms of wall time have passed (at most cap calls), then
    synchronises: a short timed loop that follows host-only work otherwise
    starts on idle GPU clocks."""
    t0 = time.perf_counter()
    for _ in range(cap):
        fn()
        if (time.perf_counter() - t0) * 1e3 >= ms:
            break
    torch.cuda.synchronize()


def bench_config1(m, torch, steps, warmup):
    """One step = rand(rng, n, u32) then rand(rng, U01OO<double>, n, f64),
    each from a fresh Philox4x32 with the configuration key, through the C
    ABI (the calls a C++ caller makes; ctypes arguments prepared once).
    The timed region holds only the steps; per-kernel durations come from a
    separate back-to-back loop of each fill between two HIP events on the
    launch stream."""
    from mckl_amd import _lib
    n = 1 << 24
    u32 = torch.empty(n, dtype=torch.int32, device="cuda")
    f64 = torch.empty(n, dtype=torch.float64, device="cuda")
    lib = _lib.lib
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    key = (ctypes.c_uint64 * 4)(KEY1[0], KEY1[1], 0, 0)
    eng = _lib.Engine()
    pe = ctypes.byref(eng)
    pu = ctypes.c_void_p(u32.data_ptr())
    pf = ctypes.c_void_p(f64.data_ptr())
    cn = ctypes.c_size_t(n)

    def raw():
        lib.mckl_hip_engine_init(pe, _lib.PHILOX4X32, 32, key)
        lib.mckl_hip_rand(pe, cn, pu, st)

    def u01():
        lib.mckl_hip_engine_init(pe, _lib.PHILOX4X32, 32, key)
        lib.mckl_hip_rand_u01_f64(pe, _lib.U01_OO, cn, pf, st)

    for _ in range(warmup):
        raw()
        u01()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        raw()
        u01()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the result of the timed steps is the configuration's stream
    rng = m.Philox4x32(key=KEY1)
    chk = torch.empty(n, dtype=torch.int32, device="cuda")
    m.rand(rng, n, chk)
    assert torch.equal(chk, u32)

    def kernel_ms(fn, reps=50):
        fn()
        torch.cuda.synchronize()
        a, b = _events()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    return {
        "n": n, "wall_s": wall, "bytes_per_step": 12 * n,
        "raw_ms": kernel_ms(raw), "u01_ms": kernel_ms(u01),
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


# VALU lane instructions per Normal sample of normal_kernel, from its
# rocprofv3 counters (SQ_INSTS_VALU x 64 / samples per dispatch), and the
# VALU roof: one wave instruction per cycle per CU, 256 CUs at 2.4 GHz
NORMAL_VALU_PER_SAMPLE = {("threefry4x64", "f32"): 139.1,
                          ("threefry4x64", "f64"): 333.6,
                          ("ars", "f32"): 156.3, ("ars", "f64"): 343.9}
VALU_PEAK_LANE_OPS = 256 * 2.4e9 * 64


def bench_normals(m, torch, reps=3):
    """Config-2 fills: 64 streams (keys {s,0,0,0}, s = 1..64) of 2^22
    Normal samples each, per engine and precision; samples per second, HBM
    bytes written, and the fraction of the VALU roof that bounds them (the
    profiled instructions per sample times the rate)."""
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
            vps = NORMAL_VALU_PER_SAMPLE[(ename, dname)]
            res["normal_%s_%s_2^28" % (ename, dname)] = {
                "Gsamples_per_s": round(64 * n / t / 1e9, 2),
                "GBps": round(out.element_size() * 64 * n / t / 1e9, 1),
                "ms": round(t * 1e3, 3),
                "roofline": {"bound": "valu",
                             "valu_lane_instr_per_sample": vps,
                             "frac": round(64 * n / t * vps /
                                           VALU_PEAK_LANE_OPS, 3),
                             "hbm_frac": round(out.element_size() * 64 * n /
                                               t / 1e9 / HBM_PEAK_GBS, 3)}}
            del out
    return res


def bench_smc(m, torch):
    res = bench_smc_plain(m, torch)
    res.update(bench_smc_generic(m, torch))
    res.update(bench_smc_batched(m, torch))
    return res


def bench_smc_plain(m, torch):
    from mckl_amd import workloads as wl
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
    return res


def bench_smc_generic(m, torch):
    """The plain configuration (LG (0.9, 1, 0.5), N = 2^20, T = 100,
    systematic when ESS < N/2) as a user would write it for the reference's
    SMCSampler<T>: evaluation objects of torch operations on the device
    state (FP64), the library's Normal fill, Weight, ResampleEval and
    select, the ESS read to the host once per iteration."""
    from mckl_amd import workloads as wl
    y = wl.lg_observations(100, 7)
    N, T = 1 << 20, 100
    a, sx, sy = 0.9, 1.0, 0.5
    c = -math.log(sy) - 0.5 * math.log(2 * math.pi)

    def run(seed):
        s = m.GenericSMCSampler(N, 1, torch.float64, seed=seed)
        eng = m.Philox4x32(seed + 1)
        z = torch.empty(N, dtype=torch.float64, device="cuda")
        dist = m.NormalDistribution(0.0, 1.0, "float64")
        lnc = []

        def sel(t, p):
            if t == 0:
                m.rand(eng, dist, N, z)
                p.state[:, 0] = z
            m.rand(eng, dist, N, z)
            x = p.state[:, 0]
            x.mul_(a).add_(z * sx)
            d = (float(y[t]) - x) / sy
            p.weight().add_log(d * d * -0.5 + c)
            lnc.append(p.weight().stats()[1:2].clone())

        s.selection(sel)
        s.resample(m.SYSTEMATIC)
        s.resample_threshold(0.5)
        s.iterate(T)
        return s, lnc

    run(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s, lnc = run(2)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ll = float(torch.cat(lnc).sum().item())
    res = int((s.ess_history() < 0.5 * N).sum())
    return {"smc_generic_lg_N2^20_T100": {
        "particle_steps_per_s": N * T / t, "ms": round(t * 1e3, 3),
        "loglik": ll, "resampling_steps": res,
        "note": "SMCSampler<T> with torch evaluation objects, FP64 state; "
                "host clock around the whole run (one ESS read per step)"}}


def bench_weight(m, torch):
    """Weight::add_log (core/weight.hpp:143-148, 198-227) on the device at
    N = 2^20 and 2^24: W_i = exp(log W_i + v_i - max) / sum, ESS, log of
    the normaliser.  Algorithmic bytes: v and W read, W written (24 B per
    weight); the three passes (log + max, exp + sums, scale) move 48."""
    out = {}
    for lg in (20, 24):
        N = 1 << lg
        g = torch.Generator(device="cuda").manual_seed(5)
        w = m.Weight(N)
        v = torch.randn(N, dtype=torch.float64, device="cuda", generator=g)
        _warm(torch, lambda: w.add_log(v))
        a, b = _events()
        reps = 50
        a.record()
        for _ in range(reps):
            w.add_log(v)
        b.record()
        torch.cuda.synchronize()
        us = a.elapsed_time(b) / reps * 1e3
        gbs = 24.0 * N / (us * 1e-6) / 1e9
        out["weight_add_log_N2^%d" % lg] = {
            "us": round(us, 1), "GBps_algorithmic": round(gbs, 1),
            "frac_of_hbm": round(gbs / 8000.0, 3)}
    return out


def bench_resample(m, torch):
    """ResampleAlgorithm + rep -> index on the device (mckl_hip_resample) at
    N = M = 2^20 on normalised weights exp(N(0, 1)) (a typical SMC step):
    microseconds per call and particles per second, per scheme, and the
    count of boundary cases among the first call's uniforms."""
    N = 1 << 20
    g = torch.Generator(device="cuda").manual_seed(3)
    w = torch.randn(N, dtype=torch.float64, device="cuda", generator=g).exp()
    w /= w.sum()
    rep = torch.empty(N, dtype=torch.int32, device="cuda")
    idx = torch.empty(N, dtype=torch.int32, device="cuda")
    out = {}
    for name, alg, scheme in (
            ("multinomial", m.ResampleMultinomial(), m.MULTINOMIAL),
            ("stratified", m.ResampleStratified(), m.STRATIFIED),
            ("systematic", m.ResampleSystematic(), m.SYSTEMATIC),
            ("residual_systematic", m.ResampleResidualSystematic(),
             m.RESIDUAL_SYSTEMATIC)):
        rng = m.Philox4x32(9)
        _warm(torch, lambda: alg(N, N, rng, w, rep, idx))
        a, b = _events()
        reps = 50
        a.record()
        for _ in range(reps):
            alg(N, N, rng, w, rep, idx)
        b.record()
        torch.cuda.synchronize()
        us = a.elapsed_time(b) / reps * 1e3
        out[name] = {"us": round(us, 1), "particles_per_s": N / (us * 1e-6)}
        out[name]["boundary_cases"] = _boundary_cases(m, torch, scheme, w, N)
    return {"resample_N2^20": out}


def _boundary_cases(m, torch, scheme, w, M):
    """Uniforms of the first call (a fresh Philox4x32(9)) that lie within
    1e-15, and within N eps, relative of a cumulative-weight boundary: the
    only places where a resampled ancestor depends on the rounding of the
    running weight sum, so where the reference's own answer is not fixed by
    the weights.  Boundaries are torch.cumsum in FP64 of the weights (the
    residual schemes: of the residual weights, with M - sum floor(M w)
    uniforms)."""
    if scheme >= m.RESIDUAL:
        mw = w * M
        fl = torch.floor(mw)
        r = mw - fl
        w = r / r.sum()
        M = M - int(fl.sum().item())
        seq = scheme - m.RESIDUAL if scheme > m.RESIDUAL else m.MULTINOMIAL
    else:
        seq = scheme
    N = w.numel()
    if M <= 0:
        return {"uniforms": 0, "within_1e-15": 0, "within_N_eps": 0}
    u = torch.empty(M, dtype=torch.float64, device="cuda")
    m.u01_rand_sequence(seq, m.Philox4x32(9), M, u)
    cw = torch.cumsum(w, 0)
    k = torch.searchsorted(cw, u).clamp(max=N - 1)
    hi = cw[k]
    lo = cw[(k - 1).clamp(min=0)]
    d = (hi - u).abs() / hi
    dl = torch.where(k > 0, (u - lo).abs() / lo.clamp(min=1e-300),
                     torch.full_like(u, 1.0))
    d = torch.minimum(d, dl)
    eps = float(torch.finfo(torch.float64).eps)
    return {"uniforms": M, "within_1e-15": int((d <= 1e-15).sum().item()),
            "within_N_eps": int((d <= N * eps).sum().item())}


def bench_pf_example(m, torch):
    """The reference's own particle-filter example (example/algorithm/
    include/algorithm_pf.hpp: 4-D state, Student-t weights through log1p,
    Normal<double> moves, Stratified resampling at 0.5) at its own N =
    100000 over its 100 observations, on the device API
    (tests/cpp/algorithm_pf, a separate process): the example's timed
    second run (clear(), then iterate), host clock around iterate."""
    prog = os.path.join(ROOT, "tests", "cpp", "algorithm_pf")
    if not os.path.exists(prog):
        return {"pf_example_N1e5_T100": "tests/cpp/algorithm_pf not built"}
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run(
            [prog, os.path.join(ROOT, "tests", "golden", "algorithm_pf.data"),
             "100000", "1", "1000", os.path.join(d, "pf.bin"), "time"],
            capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError("algorithm_pf failed: " + r.stderr[-500:])
    sec = float(r.stdout.split("iterate_seconds")[1].split()[0])
    res = {"pf_example_N1e5_T100": {
        "seconds": round(sec, 4),
        "particle_steps_per_s": 1e5 * 100 / sec,
        "note": "generic SMCSampler<double> (ColMajor 4-column state) with "
                "device functors for the selection arithmetic; one ESS read "
                "to the host per step"}}
    res.update(_pf_fused(torch))
    return res


def _pf_fused(torch):
    """The same model as a four-float device model on the fused plain
    sampler (tests/cpp/fused_model_test.hip, PF4: FP32 states, the
    samplers' own noise, stratified resampling below ESS N/2), one
    captured 100-step run on the event clock."""
    import ctypes
    import numpy as np
    lib_path = os.path.join(ROOT, "tests", "cpp", "libfused_model_test.so")
    if not os.path.exists(lib_path):
        return {"pf_example_fused_T100": "libfused_model_test.so not built"}
    fl = ctypes.CDLL(lib_path)
    fl.fused_last_error.restype = ctypes.c_char_p
    xy = np.ascontiguousarray(np.loadtxt(
        os.path.join(ROOT, "tests", "golden", "algorithm_pf.data")).ravel())
    T = len(xy) // 2
    out = {}
    for N in (100000, 1 << 20):
        ll = ctypes.c_double()
        ess = np.zeros(T)
        res = np.zeros(T, np.int32)
        post = np.zeros(2)
        ms = ctypes.c_double()
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        rc = fl.fused_pf_run(P(xy), ctypes.c_size_t(T), ctypes.c_size_t(N),
                             ctypes.c_uint64(5), ctypes.byref(ll), P(ess),
                             P(res), P(post), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(fl.fused_last_error().decode())
        out["pf_example_fused_N%d_T100" % N] = {
            "ms": round(ms.value, 3),
            "particle_steps_per_s": N * T / (ms.value * 1e-3),
            "resampling_steps": int(res.sum()),
            "posterior_mean_p_last": float(post[0])}
    return out


def bench_smc_batched_run(m, torch):
    """The batched configuration alone: one warm-up and one timed graph run
    of the 4096 SV filters (no per-filter comparison), for counter runs."""
    from mckl_amd import workloads as wl
    import numpy as np
    theta, Nf, yb, seeds = wl.sv_batch(4096, 50)
    bt = m.SMCBatch(m.MODEL_SV, Nf, theta, yb, seeds, 0.5, m.STRATIFIED)
    bt.run(use_graph=True)
    torch.cuda.synchronize()
    a, b = _events()
    a.record()
    bt.run(use_graph=True)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    ps = float(Nf.astype(np.float64).sum()) * 50
    return {"smc_batched_sv_4096_T50_run": {
        "ms": round(ms, 3), "particle_steps_per_s": ps / (ms * 1e-3)}}


def bench_smc_batched(m, torch):
    from mckl_amd import workloads as wl
    import numpy as np
    res = {}
    # batched SV filters
    theta, Nf, yb, seeds = wl.sv_batch(4096, 50)
    T = 50
    ps = float(Nf.astype(np.float64).sum()) * T
    bt = m.SMCBatch(m.MODEL_SV, Nf, theta, yb, seeds, 0.5, m.STRATIFIED)
    # the library's autotuner (its candidates checked bit-identical); the
    # chosen configuration is applied to the handle
    tune, tune_ms = bt.autotune()
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
    lp = np.array([sm.log_likelihood() for sm in samplers])
    res["smc_batched_sv_4096_T50"] = {
        "particle_steps": ps,
        "batched_particle_steps_per_s": ps / tb,
        "per_filter_particle_steps_per_s": ps / tp,
        "speedup_vs_per_filter": tp / tb,
        "batched_ms": round(tb * 1e3, 3), "per_filter_ms": round(tp * 1e3, 1),
        "tune": dict(tune, autotune_ms=round(tune_ms, 3)),
        "bit_identical_all_4096": bool(np.array_equal(lb, lp)),
    }
    return res


def bench_grad(m, torch):
    """Gradient configuration: SV (0.95, 0.2, 0.7), T = 500 observations
    from seed 11, N = 2^16, 256 jittered parameter points, checkpoints every
    50 steps; central-difference check on two points with the ancestor tape
    (the Kalman comparison is a test, tests/test_gpu_grad.py)."""
    from mckl_amd import workloads as wl
    import numpy as np
    res = {}
    y = wl.sv_observations(500, 11)
    th = wl.sv_grad_points(256)
    N, T, nb = 1 << 16, 500, 256
    g = m.SMCGradient(m.MODEL_SV, nb, N, y, np.arange(1, nb + 1), 0.5,
                      m.STRATIFIED, K=50)
    g.run(th)
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        ll, gr = g.run(th)
    t = (time.perf_counter() - t0) / reps
    del g
    ps = float(nb) * N * T
    # FD check: two points, ancestor tape, central differences h = 1e-4
    gt = m.SMCGradient(m.MODEL_SV, 2, N, y, [1, 2], 0.5, m.STRATIFIED, K=50,
                       tape=True)
    _, g2 = gt.run(th[:2])

    def fd(h):
        out = np.zeros((2, 3))
        for j in range(3):
            tp = th[:2].copy()
            tm = th[:2].copy()
            tp[:, j] += h
            tm[:, j] -= h
            out[:, j] = (gt.replay(tp) - gt.replay(tm)) / (2 * h)
        return out

    f1 = fd(1e-4)
    f2 = fd(5e-5)
    rich = (4 * f2 - f1) / 3
    del gt
    res["smc_grad_sv_256x2^16_T500"] = {
        "ms_per_gradient_batch": round(t * 1e3, 2),
        "forward_particle_steps_per_s": ps / t,
        "note": "forward + checkpoint recompute + reverse sweep; FP64 states",
        "max_rel_err_vs_fd_h1e-4": float(np.max(np.abs(f1 - g2) /
                                              np.abs(g2))),
        "max_rel_err_vs_richardson_fd": float(np.max(np.abs(rich - g2) /
                                                   np.abs(g2))),
        "mean_grad": gr.mean(0).tolist(),
    }
    res.update(bench_grad_kalman(m))
    return res


def bench_grad_kalman(m):
    """The adjoint's mean gradient on the linear-Gaussian model against the
    Kalman FP64 gradient (values in tests/golden/kalman_lg.json, written by
    tests/golden/make_kalman.py from the oracle's Kalman filter): the
    relative error of the mean and its standard error, per parameter, at
    N = 2^16 with 256 seeds on the T = 100 observations at (0.9, 1, 0.5),
    and with 16384 seeds on their first 10 at (0.7, 1.3, 0.7)."""
    import numpy as np
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "tests", "golden", "kalman_lg.json")) as f:
        ref = json.load(f)
    out = {}
    for name, nb, K in (("lg_T100", 256, 50), ("lg_T10_offset", 16384, 5)):
        r = ref[name]
        y = np.asarray(r["y"])
        ex = np.asarray(r["grad"])
        g = m.SMCGradient(m.MODEL_LG, nb, 1 << 16, y, np.arange(1, nb + 1),
                          0.5, m.SYSTEMATIC, K=K)
        t0 = time.perf_counter()
        _, gr = g.run(r["theta"])
        t = time.perf_counter() - t0
        del g
        mean = gr.mean(0)
        se = gr.std(0, ddof=1) / np.sqrt(nb)
        out["smc_grad_vs_kalman_%s" % name] = {
            "theta": r["theta"], "T": r["T"], "N": 1 << 16, "seeds": nb,
            "kalman_grad": ex.tolist(), "mean_grad": mean.tolist(),
            "rel_err": (np.abs(mean - ex) / np.abs(ex)).tolist(),
            "rel_se": (se / np.abs(ex)).tolist(),
            "ms": round(t * 1e3, 1),
        }
    return out


def bench_resample_index(m, torch, N=1 << 20, R=256, reps=5):
    """ResampleIndex::read_index_matrix over R = 256 iterations of N = 2^20
    particles, both layouts, for a genealogy of multinomial resampling
    (random parents) and one of systematic resampling of near-equal weights
    (parents mostly the particle itself, as in a well-mixed sampler).
    Algorithmic bytes: 4 written + 4 read per matrix entry."""
    res = {}
    g = torch.Generator(device="cpu").manual_seed(3)
    rng = m.Philox4x32(3)
    rep = torch.empty(N, dtype=torch.int32, device="cuda")
    for kind, alg, spread in (("multinomial", m.ResampleMultinomial(), 1.0),
                              ("systematic_flat", m.ResampleSystematic(),
                               0.05)):
        ri = m.ResampleIndex()
        for _ in range(R):
            w = 1.0 + spread * torch.rand(N, generator=g, dtype=torch.float64)
            w = (w / w.sum()).cuda()
            idx = torch.empty(N, dtype=torch.int32, device="cuda")
            alg(N, N, rng, w, rep, idx)
            ri.push_back(N, idx)
        out = torch.empty(N * R, dtype=torch.int32, device="cuda")
        for layout, lname in ((m.RowMajor, "row"), (m.ColMajor, "col")):
            ri.read_index_matrix(layout, out)
            a, b = _events()
            a.record()
            for _ in range(reps):
                ri.read_index_matrix(layout, out)
            b.record()
            b.synchronize()
            t = a.elapsed_time(b) / reps / 1e3
            res["resample_index_%s_%s_GBps" % (kind, lname)] = round(
                8.0 * N * R / t / 1e9, 1)
            res["resample_index_%s_%s_ms" % (kind, lname)] = round(t * 1e3, 3)
        del ri, out
    res["resample_index_config"] = "N=2^20, R=256, u32 entries"
    return res


def _traffic():
    """HBM bytes per launch of the U01OO<double> fill (FETCH_SIZE +
    WRITE_SIZE, each scaled by the same counter on a copy of known size),
    from the rocprofv3 counter summary kept with the profiles; None if that
    summary is absent."""
    path = os.path.join(ROOT, "profiles", "u01f64_fill_config1_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return int(float(line.split("=")[1]))
    except OSError:
        pass
    return None


def _guarded(fn, *a):
    try:
        return fn(*a)
    except Exception as exc:  # a CPU side figure must not sink the GPU line
        return {"error": repr(exc)}


def cpu_reference_points(threads):
    """CPU figures beside the secondary GPU legs, on this host: the
    reference compiled from its headers (fills, Normal streams, the
    ResampleIndex trace-back), in its pinned build (-O2, SSE2) and its speed
    build (-O3, x86-64-v4 + AES-NI) where the host has those features, on one
    thread and on `threads`; and the oracle port for the SMC samplers (the
    reference has no counterpart of the fused samplers).  Bounded samples:
    the config-1 step, 2^22 normals per stream-set on one thread and the full
    64 x 2^22 on `threads`, the plain configuration over its first 10 steps,
    the first 64 filters of the batched configuration, the trace-back of 64
    iterations of 2^20 particles."""
    import numpy as np
    from mckl_amd import workloads as wl
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    res = {"threads": threads, "nproc": os.cpu_count(), "lscpu": host_cpu(),
           "build": "oracle port: gcc -O2 -ffp-contract=off -fopenmp (its "
                    "pinned flags); reference pinned: g++ -O2 "
                    "-ffp-contract=off, no -march; reference speed: g++ -O3 "
                    "-march=x86-64-v4 -maes (%s)" % (
                        "loaded" if orc.REF_V4 is not None else
                        "not loaded: host lacks AVX-512/AES-NI or not built")}
    builds = [(nm, lib) for nm, lib in (("pinned", orc.REF),
                                         ("speed", orc.REF_V4))
              if lib is not None]
    n = 1 << 24
    u32 = np.empty(n, np.uint32)
    f64 = np.empty(n, np.float64)
    key = (ctypes.c_uint64 * 4)(*(KEY1 + [0, 0]))
    for bname, lib in builds:
        for nt in (1, threads):
            t = lib.ref_time_philox_u01_mt(key, ctypes.c_size_t(n), 5, nt,
                                           orc.P(u32), orc.P(f64))
            res["config1_step_reference_%s_%dthreads_GBps" % (bname, nt)] = \
                round(12 * n / t / 1e9, 3)
    del u32, f64
    for bname, lib in builds:
        for ename, kind, dname, dbl, mean, sd in (
                ("threefry4x64", orc.THREEFRY4X64, "f32", 0, 0.0, 1.0),
                ("threefry4x64", orc.THREEFRY4X64, "f64", 1, 2.5, 0.3),
                ("ars", orc.ARS, "f32", 0, 0.0, 1.0),
                ("ars", orc.ARS, "f64", 1, 2.5, 0.3)):
            for nt, ns in ((1, 1), (threads, 64)):
                n = 1 << 22
                out = np.empty(ns * n, np.float64 if dbl else np.float32)
                t = lib.ref_time_normal_streams(kind, dbl, ns, n, mean, sd, nt,
                                                orc.P(out))
                del out
                res["normal_%s_%s_reference_%s_%dthreads_Gsamples_per_s" % (
                    ename, dname, bname, nt)] = round(ns * n / t / 1e9, 4)
    y = wl.lg_observations(100, 7)[:10]
    N = 1 << 20
    for nt in (1, threads):
        t0 = time.perf_counter()
        orc.smc_run(1, (0.9, 1.0, 0.5), y, N, 1, 0.5, orc.SYSTEMATIC,
                    nthreads=nt)
        t = time.perf_counter() - t0
        res["smc_plain_lg_port_%dthreads_particle_steps_per_s" % nt] = \
            N * len(y) / t
    theta, Nf, yb, seeds = wl.sv_batch(4096, 50)
    ps = 0.0
    t0 = time.perf_counter()
    for f in range(64):
        orc.smc_run(2, theta[f], yb[f], int(Nf[f]), int(seeds[f]), 0.5,
                    orc.STRATIFIED, nthreads=threads)
        ps += float(Nf[f]) * 50
    t = time.perf_counter() - t0
    res["smc_batched_sv_port_%dthreads_particle_steps_per_s" % threads] = \
        ps / t
    pf_lib = os.path.join(orc.HERE, "_ref", "libmckl_ref_pf.so")
    if os.path.exists(pf_lib):
        # the reference's particle-filter example itself, BackendSEQ, one
        # thread, its own N = 100000 over its 100 observations (sampler
        # construction and the data read included)
        lib = ctypes.CDLL(pf_lib, mode=os.RTLD_LAZY)
        N = 100000
        Tn = ctypes.c_size_t()
        ess = np.zeros(1000)
        size = np.zeros(1000, np.uint64)
        st = np.zeros(4 * N)
        wv = np.zeros(N)
        keys = np.zeros(4, np.uint64)
        t0 = time.perf_counter()
        rc = lib.ref_pf(ctypes.c_size_t(N), ctypes.c_uint64(1),
                        os.path.join(ROOT, "tests", "golden").encode(),
                        ctypes.c_size_t(1000), ctypes.byref(Tn), orc.P(ess),
                        orc.P(size), orc.P(st), orc.P(wv), orc.P(keys))
        t = time.perf_counter() - t0
        if rc == 0:
            res["pf_example_N1e5_T100_reference_seq_1thread_s"] = round(t, 3)
    if orc.REF is not None:
        # the reference's read_index_matrix (column-major) over 64 iterations
        # of 2^20 uniformly random parents, 8 algorithmic bytes per entry
        N, R = 1 << 20, 64
        g = np.random.default_rng(5)
        orc.REF.ref_ri_new()
        for _ in range(R):
            v = g.integers(0, N, N).astype(np.uint32)
            orc.REF.ref_ri_push(ctypes.c_size_t(N), orc.P(v))
        buf = np.zeros(N * R, np.uint32)
        t0 = time.perf_counter()
        orc.REF.ref_ri_matrix(102, ctypes.c_size_t(R - 1), ctypes.c_size_t(0),
                              orc.P(buf))
        t = time.perf_counter() - t0
        res["resample_index_col_reference_1thread_GBps"] = round(
            8.0 * N * R / t / 1e9, 3)
        orc.REF.ref_ri_new()
    return res


def host_threads():
    """CPU threads this process may use: its affinity mask, at most 16 (the
    host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_cpu():
    """lscpu's description of the host CPU (model name, family/model/
    stepping, topology, caches, SIMD flags)."""
    info = {}
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True,
                             timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "Flags":
                fl = set(v.split())
                info["simd_flags"] = " ".join(
                    f for f in ("sse4_2", "avx2", "fma", "bmi2", "aes", "vaes",
                                "avx512f", "avx512bw", "avx512vl",
                                "avx512dq", "avx512_vnni") if f in fl)
            elif k in ("Model name", "Vendor ID", "CPU family", "Model",
                       "Stepping", "CPU(s)", "Thread(s) per core",
                       "Core(s) per socket", "Socket(s)", "L2 cache",
                       "L3 cache", "CPU max MHz", "Hypervisor vendor",
                       "NUMA node(s)"):
                info[k] = v
    except (OSError, ValueError):
        pass
    # a hypervisor may shorten the model name (to "AMD EPYC"); the family
    # number still names the core generation
    if info.get("Vendor ID") == "AuthenticAMD":
        fam = info.get("CPU family", "")
        info["core_generation"] = {"23": "Zen 1/2", "25": "Zen 3/4",
                                   "26": "Zen 5"}.get(fam, "family " + fam)
    return info


def cpu_baseline(seconds=10.0):
    """The reference's own bulk fills, compiled from its headers: its speed
    build (-O3, x86-64-v4 + AES-NI, so its AVX-512 Philox and U01 paths) where
    the host has those features, else its pinned build, on every host thread
    allotted (at most 16), the step split by counter range (each thread's
    engines moved to its range by discard; the outputs equal the
    single-thread ones).  Else the oracle port on one thread.  Sample: the
    full config-1 step, repeated for about `seconds` (at least 10 runs),
    median per step."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    n = 1 << 24
    u32 = np.empty(n, np.uint32)
    f64 = np.empty(n, np.float64)
    lib = orc.REF_V4 if orc.REF_V4 is not None else orc.REF
    threads = host_threads()
    if lib is not None:
        key = (ctypes.c_uint64 * 4)(*(KEY1 + [0, 0]))
        t1 = lib.ref_time_philox_u01_mt(key, ctypes.c_size_t(n), 3, threads,
                                        orc.P(u32), orc.P(f64))
        steps = int(min(20000, max(10, seconds / max(t1, 1e-6))))
        t = lib.ref_time_philox_u01_mt(key, ctypes.c_size_t(n), steps,
                                       threads, orc.P(u32), orc.P(f64))
        build = ("g++ -O3 -march=x86-64-v4 -maes (AVX-512, AES-NI)"
                 if lib is orc.REF_V4 else "g++ -O2 -ffp-contract=off")
        return {"value": round(12 * n / t / 1e9, 3), "unit": "GB/s",
                "cores": threads, "kind": "reference",
                "sample": "full step: 2^24 u32 + 2^24 U01OO doubles, "
                          "Philox4x32, reference built %s, %d threads split "
                          "by counter range, median of %d runs"
                          % (build, threads, steps)}
    ts = []
    for _ in range(10):
        t0 = time.perf_counter()
        e = orc.engine(orc.PHILOX4X32, KEY1)
        orc.O.mo_rand_u32(ctypes.byref(e), ctypes.c_size_t(n), orc.P(u32))
        e = orc.engine(orc.PHILOX4X32, KEY1)
        orc.O.mo_rand_u01_f64(ctypes.byref(e), 3, ctypes.c_size_t(n),
                              orc.P(f64))
        ts.append(time.perf_counter() - t0)
    t = sorted(ts)[len(ts) // 2]
    return {"value": round(12 * n / t / 1e9, 3), "unit": "GB/s", "cores": 1,
            "kind": "port",
            "sample": "full step: 2^24 u32 + 2^24 U01OO doubles, Philox4x32, "
                      "oracle port, median of 10 runs, single thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--only", choices=["fills", "saturation", "normals", "smc",
                                       "smc_plain", "smc_generic",
                                       "smc_batched", "smc_batched_run",
                                       "grad", "pf_example",
                                       "resample_index", "resample",
                                       "weight"],
                    help="run one secondary leg alone and print its JSON "
                         "(for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the fills and samplers shard by "
              "counter range / filter with no communication, so no multi-GPU "
              "run is made; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    import mckl_amd as m

    if args.only == "fills":
        # the config-1 fills plus a device copy of known size (2^24 doubles
        # read and written), the scale for the HBM byte counters
        c1 = bench_config1(m, torch, args.steps, args.warmup)
        src = torch.ones(1 << 24, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({k: v for k, v in c1.items()}), flush=True)
        return
    if args.only:
        leg = {"saturation": bench_saturation, "normals": bench_normals,
               "smc": bench_smc, "smc_plain": bench_smc_plain,
               "smc_generic": bench_smc_generic,
               "smc_batched": bench_smc_batched,
               "smc_batched_run": bench_smc_batched_run,
               "grad": bench_grad,
               "pf_example": bench_pf_example,
               "weight": bench_weight,
               "resample_index": bench_resample_index,
               "resample": bench_resample}[args.only]
        print(json.dumps(leg(m, torch)), flush=True)
        return
    c1 = bench_config1(m, torch, args.steps, args.warmup)
    value = c1["bytes_per_step"] * args.steps / c1["wall_s"] / 1e9
    u01_gbs = 8 * c1["n"] / (c1["u01_ms"] / 1e3) / 1e9
    raw_gbs = 4 * c1["n"] / (c1["raw_ms"] / 1e3) / 1e9
    secondary = {"raw_u32_fill_GBps": round(raw_gbs, 1),
                 "raw_u32_fill_ms": round(c1["raw_ms"], 4),
                 "u01oo_f64_fill_GBps": round(u01_gbs, 1),
                 "u01oo_f64_fill_ms": round(c1["u01_ms"], 4)}
    if not args.no_secondary:
        threads = host_threads()
        for name, leg in (("saturation", bench_saturation),
                          ("normals", bench_normals), ("smc", bench_smc),
                          ("grad", bench_grad),
                          ("pf_example", bench_pf_example),
                          ("resample_index", bench_resample_index),
                          ("resample", bench_resample),
                          ("weight", bench_weight),
                          ("cpu", lambda m_, t_: {
                              "cpu": _guarded(cpu_reference_points,
                                              threads)})):
            t0 = time.perf_counter()
            secondary.update(leg(m, torch))
            print("bench: %s leg done in %.1f s" % (name,
                  time.perf_counter() - t0), file=sys.stderr, flush=True)
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
            "traffic": _traffic(),
        },
        "cpu_baseline": cpu_baseline(),
        "secondary": secondary,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
