// smc_plain.hpp -- host side of the plain fused SMC sampler, as a template
// on the model (the model concept is in smc_tiles.hpp): device buffers, the
// launch sequence of a T-step run captured once into a HIP graph, results.
// The C ABI (smc.hip) instantiates it on the built-in models; a program
// compiled with hipcc instantiates it on its own model through
// mckl_amd/include/mckl_hip_fused.hpp.  Every function returns an
// MCKL_HIP_* status (set_error records the message).
#pragma once

#include "common.hpp"
#include "smc_kernels.hpp"

#include <cmath>
#include <cstring>

namespace mckl_hip {

inline int check_common(size_t N, size_t T, double threshold, int scheme)
{
    if (N == 0 || N >= (size_t)0x7FFFFFFF)
        return set_error(MCKL_HIP_EINVAL, "particle count %zu out of range", N);
    if (T == 0)
        return set_error(MCKL_HIP_EINVAL, "T must be > 0");
    if (scheme != MCKL_HIP_MULTINOMIAL && scheme != MCKL_HIP_STRATIFIED &&
        scheme != MCKL_HIP_SYSTEMATIC &&
        !is_residual(scheme))
        return set_error(MCKL_HIP_EINVAL,
            "resampling scheme %d is not provided by this sampler", scheme);
    if (std::isnan(threshold))
        return set_error(MCKL_HIP_EINVAL, "threshold is NaN");
    return MCKL_HIP_OK;
}

template <class M>
struct PlainSampler {
    SmcArgs A{};
    M model{};
    size_t T = 0;
    float *x[2] = {nullptr, nullptr};
    float *lw = nullptr, *tmax = nullptr;
    uint32_t *idx = nullptr;
    double *ts = nullptr, *ts2 = nullptr;
    RsBuf B{};
    SmcState *st = nullptr;
    double *ess = nullptr;
    int32_t *res = nullptr;
    int kmom = 0; // estimated moments
    double *eph = nullptr, *epl = nullptr, *est = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // recorded on the caller's stream after each run: result calls wait for
    // it, not for the whole device, and read through a non-blocking stream
    // of their own (the null stream would wait for every blocking stream)
    hipEvent_t done = nullptr;
    hipStream_t rs = nullptr;
};

template <class M>
inline void plain_drop_graph(PlainSampler<M> *s)
{
    if (s->exec)
        (void)hipGraphExecDestroy(s->exec);
    if (s->graph)
        (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
}

// frees everything plain_init allocated (the object itself stays)
template <class M>
inline void plain_release(PlainSampler<M> *s)
{
    plain_drop_graph(s);
    if (s->done)
        (void)hipEventDestroy(s->done);
    if (s->rs)
        (void)hipStreamDestroy(s->rs);
    s->done = nullptr;
    s->rs = nullptr;
    auto drop = [](auto *&p) {
        if (p)
            (void)hipFree(p);
        p = nullptr;
    };
    drop(s->x[0]);
    drop(s->x[1]);
    drop(s->lw);
    drop(s->tmax);
    drop(s->idx);
    drop(s->ts);
    drop(s->ts2);
    drop(s->B.tdh);
    drop(s->B.tdl);
    drop(s->B.te);
    drop(s->B.tz);
    drop(s->B.cnt);
    drop(s->B.tmh);
    drop(s->B.tml);
    drop(s->B.u);
    drop(s->B.uh);
    drop(s->B.ul);
    drop(s->B.ti);
    drop(s->B.trh);
    drop(s->B.trl);
    drop(s->B.src);
    drop(s->B.lbw);
    drop(s->B.tpre);
    drop(s->st);
    drop(s->ess);
    drop(s->res);
    drop(s->eph);
    drop(s->epl);
    drop(s->est);
    s->kmom = 0;
}

// Sizes, RNG key, threshold and scheme; buffers for N particles and T
// steps.  On failure everything allocated is released.
template <class M>
inline int plain_init(PlainSampler<M> *s, const M &model, size_t N, size_t T,
    uint64_t seed, double threshold, int scheme)
{
    int rc = check_common(N, T, threshold, scheme);
    if (rc)
        return rc;
    s->model = model;
    s->T = T;
    s->A.k0 = (uint32_t)seed;
    s->A.k1 = (uint32_t)(seed >> 32);
    s->A.N = (uint32_t)N;
    s->A.ntiles = (uint32_t)ceil_div(N, kTile);
    s->A.tauN = threshold * (double)N;
    s->A.scheme = scheme;
    s->A.T = (uint32_t)T;
    const size_t nt = s->A.ntiles;
    const size_t Np = nt * kTile; // padded: vector accesses stay in bounds
#define MCKL_ALLOC(p, n)                                                      \
    do {                                                                      \
        rc = dev_alloc(&(p), (n));                                            \
        if (rc) {                                                             \
            plain_release(s);                                                 \
            return rc;                                                        \
        }                                                                     \
    } while (0)
    // states by component (model_dim of them), each a stride of Np
    MCKL_ALLOC(s->x[0], Np * (size_t)model_dim<M>::value);
    MCKL_ALLOC(s->x[1], Np * (size_t)model_dim<M>::value);
    MCKL_ALLOC(s->lw, Np);
    MCKL_ALLOC(s->idx, Np);
    MCKL_ALLOC(s->tmax, nt);
    MCKL_ALLOC(s->ts, nt);
    MCKL_ALLOC(s->ts2, nt);
    MCKL_ALLOC(s->B.tdh, nt);
    MCKL_ALLOC(s->B.tdl, nt);
    MCKL_ALLOC(s->B.te, nt);
    MCKL_ALLOC(s->B.tz, nt);
    MCKL_ALLOC(s->B.cnt, Np);
    MCKL_ALLOC(s->B.src, Np);
    MCKL_ALLOC(s->B.lbw, 2 * nt);
    if (scheme == MCKL_HIP_SYSTEMATIC || scheme == MCKL_HIP_STRATIFIED)
        MCKL_ALLOC(s->B.tpre, 2 * nt * (size_t)kGroup);
    if (is_residual(scheme)) {
        MCKL_ALLOC(s->B.ti, nt);
        MCKL_ALLOC(s->B.trh, nt);
        MCKL_ALLOC(s->B.trl, nt);
    }
    if (scheme == MCKL_HIP_MULTINOMIAL || scheme == MCKL_HIP_RESIDUAL) {
        MCKL_ALLOC(s->B.tmh, nt);
        MCKL_ALLOC(s->B.tml, nt);
        MCKL_ALLOC(s->B.u, Np);
        MCKL_ALLOC(s->B.uh, nt);
        MCKL_ALLOC(s->B.ul, nt);
    }
    MCKL_ALLOC(s->st, 1);
    MCKL_ALLOC(s->ess, T);
    MCKL_ALLOC(s->res, T);
#undef MCKL_ALLOC
    hipError_t err = hipMemset(s->st, 0, sizeof(SmcState));
    if (err == hipSuccess) // every parent a code may name stays below N
        err = hipMemset(s->B.src, 0, Np * sizeof(uint32_t));
    if (err == hipSuccess)
        err = hipEventCreateWithFlags(&s->done, hipEventDisableTiming);
    if (err == hipSuccess)
        err = hipStreamCreateWithFlags(&s->rs, hipStreamNonBlocking);
    if (err != hipSuccess) {
        rc = hip_status(err, "smc sampler set-up");
        plain_release(s);
        return rc;
    }
    return MCKL_HIP_OK;
}

template <typename... Args>
inline void launch_rs_counts(int scheme, dim3 grid, hipStream_t st,
    Args... a)
{
    switch (scheme) {
    case MCKL_HIP_MULTINOMIAL:
        k_rs_counts<MCKL_HIP_MULTINOMIAL><<<grid, kGroup, 0, st>>>(a...);
        break;
    case MCKL_HIP_STRATIFIED:
        k_rs_counts<MCKL_HIP_STRATIFIED><<<grid, kGroup, 0, st>>>(a...);
        break;
    case MCKL_HIP_SYSTEMATIC:
        k_rs_counts<MCKL_HIP_SYSTEMATIC><<<grid, kGroup, 0, st>>>(a...);
        break;
    case MCKL_HIP_RESIDUAL:
        k_rs_counts<MCKL_HIP_RESIDUAL><<<grid, kGroup, 0, st>>>(a...);
        break;
    case MCKL_HIP_RESIDUAL_STRATIFIED:
        k_rs_counts<MCKL_HIP_RESIDUAL_STRATIFIED><<<grid, kGroup, 0, st>>>(
            a...);
        break;
    default:
        k_rs_counts<MCKL_HIP_RESIDUAL_SYSTEMATIC><<<grid, kGroup, 0, st>>>(
            a...);
        break;
    }
}

// Enqueue the whole run: init, then per step the move with the tile sums,
// and the decision with the resampling counts and the rep->index transform
// (systematic, stratified: one kernel; multinomial and residual: their
// count kernels, then k_rs_index).  The resampling kernels exit at once on
// steps that do not resample.
template <class M>
inline int plain_enqueue(PlainSampler<M> *s, hipStream_t st)
{
    const dim3 grid(s->A.ntiles);
    k_init<M><<<grid, kGroup, 0, st>>>(
        s->A, s->model, s->x[0], s->lw, s->st, s->B);
    MCKL_HIP_CHECK_LAUNCH("k_init");
    for (uint32_t t = 1; t <= s->T; ++t) {
        const float *xin = s->x[(t - 1) & 1];
        float *xout = s->x[t & 1];
        k_move<M><<<grid, kGroup, 0, st>>>(s->A, s->model, t, xin, xout,
            s->lw, s->idx, s->st, s->tmax, s->ts, s->ts2, s->B);
        if (s->kmom)
            k_est_moments<<<grid, kGroup, 0, st>>>(s->A, t, xout, s->lw,
                s->tmax, s->kmom, s->eph, s->epl);
        launch_rs_counts(s->A.scheme, grid, st, s->A, t, s->lw, s->tmax,
            s->ts, s->ts2, s->B, s->st, s->ess, s->res, s->idx);
        if (s->A.scheme == MCKL_HIP_MULTINOMIAL) {
            k_rs_uniforms<<<grid, kGroup, 0, st>>>(s->A, t, s->st, s->B);
            k_rs_mcounts<<<grid, kGroup, 0, st>>>(
                s->A, t, s->lw, s->tmax, s->st, s->B);
        } else if (is_residual(s->A.scheme)) {
            if (s->A.scheme == MCKL_HIP_RESIDUAL) {
                k_rs_rterms<<<grid, kGroup, 0, st>>>(s->A, t, s->st, s->B);
                k_rs_uniforms<<<grid, kGroup, 0, st>>>(s->A, t, s->st, s->B);
            }
            k_rs_rcounts<<<grid, kGroup, 0, st>>>(
                s->A, t, s->lw, s->tmax, s->st, s->B);
        }
        if (s->A.scheme == MCKL_HIP_MULTINOMIAL || is_residual(s->A.scheme))
            k_rs_index<<<grid, kGroup, 0, st>>>(s->A, t, s->st, s->B, s->idx);
    }
    if (s->kmom)
        k_est_moments_final<<<(uint32_t)s->T, kGroup, 0, st>>>(
            s->A.ntiles, s->kmom, s->eph, s->epl, s->est);
    MCKL_HIP_CHECK_LAUNCH("smc step kernels");
    return MCKL_HIP_OK;
}

// the selection estimator of the moments x^1..x^k at every step (k = 0
// turns it off); takes effect at the next run
template <class M>
inline int plain_set_moments(PlainSampler<M> *s, int k)
{
    if (k < 0 || k > kMaxMoments)
        return set_error(MCKL_HIP_EINVAL,
            "moments: k = %d outside 0..%d", k, kMaxMoments);
    MCKL_HIP_TRY(hipEventSynchronize(s->done));
    plain_drop_graph(s);
    for (double **p : {&s->eph, &s->epl, &s->est}) {
        if (*p)
            MCKL_HIP_TRY(hipFree(*p));
        *p = nullptr;
    }
    s->kmom = 0;
    if (k == 0)
        return MCKL_HIP_OK;
    const size_t np = s->T * (size_t)(k + 1) * s->A.ntiles;
    int rc = dev_alloc(&s->eph, np);
    if (rc == MCKL_HIP_OK)
        rc = dev_alloc(&s->epl, np);
    if (rc == MCKL_HIP_OK)
        rc = dev_alloc(&s->est, s->T * (size_t)k);
    if (rc)
        return rc;
    s->kmom = k;
    return MCKL_HIP_OK;
}

template <class M>
inline int plain_estimates(PlainSampler<M> *s, double *h_est)
{
    if (!s->kmom)
        return set_error(MCKL_HIP_EINVAL,
            "no estimator: set the moments to estimate first");
    MCKL_HIP_TRY(hipEventSynchronize(s->done));
    MCKL_HIP_TRY(hipMemcpyWithStream(h_est, s->est,
        s->T * s->kmom * sizeof(double), hipMemcpyDeviceToHost, s->rs));
    return MCKL_HIP_OK;
}

// one run on `st`: the T-step launch sequence is captured once into a HIP
// graph and replayed; a run contains no host synchronisation
template <class M>
inline int plain_run(PlainSampler<M> *s, hipStream_t st)
{
    if (!s->exec) {
        hipStream_t cap;
        MCKL_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t err =
            hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int rc = MCKL_HIP_OK;
        if (err == hipSuccess) {
            rc = plain_enqueue(s, cap);
            hipGraph_t g = nullptr;
            err = hipStreamEndCapture(cap, &g);
            s->graph = g;
        }
        if (err == hipSuccess && rc == MCKL_HIP_OK)
            err = hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
        if (rc)
            return rc;
        MCKL_HIP_TRY(err);
    }
    MCKL_HIP_TRY(hipGraphLaunch(s->exec, st));
    MCKL_HIP_TRY(hipEventRecord(s->done, st));
    return MCKL_HIP_OK;
}

// log-likelihood estimate, ESS and resampling decisions of the last run
// (either pointer may be null); waits for that run only
template <class M>
inline int plain_result(PlainSampler<M> *s, double *h_loglik, double *h_ess,
    int32_t *h_resampled)
{
    MCKL_HIP_TRY(hipEventSynchronize(s->done));
    SmcState hs;
    MCKL_HIP_TRY(hipMemcpyWithStream(&hs, s->st, sizeof(hs),
        hipMemcpyDeviceToHost, s->rs));
    if (hs.err)
        return set_error(MCKL_HIP_EHIP,
            "smc run: a resampling look-back wait gave up; results invalid");
    if (h_loglik)
        *h_loglik = hs.logz;
    if (h_ess)
        MCKL_HIP_TRY(hipMemcpyWithStream(h_ess, s->ess, s->T * sizeof(double),
            hipMemcpyDeviceToHost, s->rs));
    if (h_resampled)
        MCKL_HIP_TRY(hipMemcpyWithStream(h_resampled, s->res,
            s->T * sizeof(int32_t), hipMemcpyDeviceToHost, s->rs));
    return MCKL_HIP_OK;
}

} // namespace mckl_hip
