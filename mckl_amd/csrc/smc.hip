// smc.hip -- the C ABI of the fused SMC samplers (plain, batched, the
// autotuner and the selection estimator) on the built-in linear-Gaussian
// and stochastic-volatility models; the kernels are in smc_kernels.hpp.
#include "common.hpp"
#include "smc_kernels.hpp"
#include "smc_plain.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace mckl_hip {

// -------------------------------------------------------------------------
// host side
// -------------------------------------------------------------------------
static int model_const(int model, const double *th, ModelConst &mc)
{
    if (model != 1 && model != 2)
        return set_error(MCKL_HIP_EINVAL, "unknown model %d", model);
    if (!th)
        return set_error(MCKL_HIP_EINVAL, "theta is NULL");
    if (!(th[2] > 0) || !(th[1] > 0))
        return set_error(MCKL_HIP_EINVAL, "scale parameters must be > 0");
    if (model == 2 && !(std::fabs(th[0]) < 1))
        return set_error(MCKL_HIP_EINVAL, "SV requires |phi| < 1");
    const double l2pi = std::log(2.0 * 3.141592653589793);
    mc.model = model;
    mc.a = (float)th[0];
    mc.s = (float)th[1];
    mc.cst = (float)(-std::log(th[2]) - 0.5 * l2pi);
    if (model == 1) {
        mc.sd0 = 1.0f;
        mc.inv_sy = (float)(1.0 / th[2]);
    } else {
        mc.sd0 = (float)std::sqrt(th[1] * th[1] / (1.0 - th[0] * th[0]));
        mc.inv_sy = 0.0f;
    }
    return MCKL_HIP_OK;
}

static void obs_const(int model, const double *th, const double *y, size_t T,
    float *out)
{
    for (size_t t = 0; t < T; ++t)
        out[t] = model == 1 ? (float)y[t] : (float)(y[t] * y[t] / (th[2] * th[2]));
}

} // namespace mckl_hip

using namespace mckl_hip;

// the plain sampler on the built-in models; obs holds the T observation
// constants the model reads
struct mckl_hip_smc : PlainSampler<BuiltinModel> {
    float *obs = nullptr;
};

static void smc_free(mckl_hip_smc *s)
{
    plain_release(s);
    if (s->obs)
        (void)hipFree(s->obs);
    delete s;
}

extern "C" {

int mckl_hip_smc_create(int model, const double *theta, size_t T,
    const double *h_y, size_t N, uint64_t seed, double threshold, int scheme,
    mckl_hip_smc **out)
{
    if (!out || !h_y)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    *out = nullptr;
    int rc = check_common(N, T, threshold, scheme);
    if (rc)
        return rc;
    ModelConst mc;
    rc = model_const(model, theta, mc);
    if (rc)
        return rc;
    mckl_hip_smc *s = new mckl_hip_smc();
    rc = dev_alloc(&s->obs, T);
    if (rc) {
        smc_free(s);
        return rc;
    }
    std::vector<float> ob(T);
    obs_const(model, theta, h_y, T, ob.data());
    hipError_t err = hipMemcpy(s->obs, ob.data(), T * sizeof(float),
        hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        rc = hip_status(err, "mckl_hip_smc_create");
        smc_free(s);
        return rc;
    }
    s->A.mc = mc;
    s->A.obs = s->obs;
    rc = plain_init<BuiltinModel>(s, builtin_model(s->A), N, T, seed,
        threshold, scheme);
    if (rc) {
        smc_free(s);
        return rc;
    }
    *out = s;
    return MCKL_HIP_OK;
}

int mckl_hip_smc_estimate_moments(mckl_hip_smc *s, int k)
{
    if (!s)
        return set_error(MCKL_HIP_EINVAL, "sampler is NULL");
    return plain_set_moments<BuiltinModel>(s, k);
}

int mckl_hip_smc_estimates(mckl_hip_smc *s, double *h_est)
{
    if (!s || !h_est)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (!s->kmom)
        return set_error(MCKL_HIP_EINVAL,
            "no estimator: call mckl_hip_smc_estimate_moments first");
    return plain_estimates<BuiltinModel>(s, h_est);
}

int mckl_hip_smc_run(mckl_hip_smc *s, void *stream)
{
    if (!s)
        return set_error(MCKL_HIP_EINVAL, "sampler is NULL");
    return plain_run<BuiltinModel>(s, as_stream(stream));
}

int mckl_hip_smc_result(mckl_hip_smc *s, double *h_loglik, double *h_ess,
    int32_t *h_resampled)
{
    if (!s)
        return set_error(MCKL_HIP_EINVAL, "sampler is NULL");
    return plain_result<BuiltinModel>(s, h_loglik, h_ess, h_resampled);
}

int mckl_hip_smc_destroy(mckl_hip_smc *s)
{
    if (s)
        smc_free(s);
    return MCKL_HIP_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// batched sampler host side
// ---------------------------------------------------------------------------
struct mckl_hip_smc_batch {
    size_t nfilters = 0;
    size_t T = 0;
    FilterDescT<BuiltinModel> *fd = nullptr;
    uint32_t *order = nullptr;
    uint32_t *queue = nullptr;
    float *xa = nullptr, *xb = nullptr, *lw = nullptr, *obs = nullptr;
    uint32_t *idx = nullptr, *cnt = nullptr, *E = nullptr, *zpos = nullptr;
    double *u = nullptr; // multinomial uniforms
    double *logz = nullptr;
    mckl_hip_tune tune{256, 0, 0, 256};
    // filters order[0, nlarge) run in the launch of tune.block-thread
    // workgroups (`persistent` of them), the rest beside it in one of
    // tune.small_block threads (`persistent_small`)
    uint32_t nlarge = 0;
    int persistent = 0;
    int persistent_small = 0;
    hipStream_t side = nullptr;              // the second launch's stream
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<uint32_t> nsorted;           // N in queue order
    double threshold = 0.0;
    int scheme = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    long long graph_tune = -1;
    // recorded on the caller's stream after each run, and the stream the
    // results are read through (see mckl_hip_smc)
    hipEvent_t done = nullptr;
    hipStream_t rs = nullptr;
    // host copies for set_params: descriptors, model, observations
    int model = 0;
    std::vector<FilterDescT<BuiltinModel>> fdh;
    std::vector<double> y;
};

static void batch_free(mckl_hip_smc_batch *b)
{
    if (b->side)
        (void)hipStreamDestroy(b->side);
    if (b->fork)
        (void)hipEventDestroy(b->fork);
    if (b->join)
        (void)hipEventDestroy(b->join);
    if (b->done)
        (void)hipEventDestroy(b->done);
    if (b->rs)
        (void)hipStreamDestroy(b->rs);
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    void *ptrs[] = {b->fd, b->order, b->queue, b->xa, b->xb, b->lw, b->obs,
        b->idx, b->cnt, b->E, b->zpos, b->u, b->logz};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    delete b;
}

static void launch_k_batch(int G, uint32_t grid, hipStream_t st,
    const mckl_hip_smc_batch *b, const uint32_t *order, uint32_t nf,
    uint32_t *queue)
{
    switch (G) {
        case 1:
            k_batch<1, BuiltinModel><<<grid, kGroup, 0, st>>>(b->fd, order, nf, queue,
                b->xa, b->xb, b->lw, b->idx, b->cnt, b->E, b->zpos, b->u,
                b->logz);
            break;
        case 2:
            k_batch<2, BuiltinModel><<<grid, 2 * kGroup, 0, st>>>(b->fd, order, nf, queue,
                b->xa, b->xb, b->lw, b->idx, b->cnt, b->E, b->zpos, b->u,
                b->logz);
            break;
        default:
            k_batch<4, BuiltinModel><<<grid, 4 * kGroup, 0, st>>>(b->fd, order, nf, queue,
                b->xa, b->xb, b->lw, b->idx, b->cnt, b->E, b->zpos, b->u,
                b->logz);
            break;
    }
}

// The persistent launch(es) of a run: the large filters on `st`, the others
// on the handle's side stream, forked from and joined back into `st` (both
// captured into the same graph).  Each launch takes its filters from its
// own queue counter; every workgroup exits when its queue is empty.
static int batch_launch(mckl_hip_smc_batch *b, hipStream_t st)
{
    k_zero_u32<<<1, 64, 0, st>>>(b->queue, 2);
    const uint32_t nf = (uint32_t)b->nfilters;
    const uint32_t nl = b->nlarge;
    if (nl > 0)
        launch_k_batch(std::max(1, b->tune.block / kGroup),
            (uint32_t)b->persistent, st, b, b->order, nl, b->queue);
    if (nl < nf) {
        MCKL_HIP_TRY(hipEventRecord(b->fork, st));
        MCKL_HIP_TRY(hipStreamWaitEvent(b->side, b->fork, 0));
        launch_k_batch(std::max(1, b->tune.small_block / kGroup),
            (uint32_t)b->persistent_small, b->side, b, b->order + nl, nf - nl,
            b->queue + 1);
        MCKL_HIP_TRY(hipEventRecord(b->join, b->side));
        MCKL_HIP_TRY(hipStreamWaitEvent(st, b->join, 0));
    }
    MCKL_HIP_CHECK_LAUNCH("k_batch");
    return MCKL_HIP_OK;
}

// resident workgroups per CU of k_batch<G>, capped by waves_per_eu
static int batch_per_cu(int G, int waves_per_eu, int &per_cu)
{
    int occ = 0;
    hipError_t err;
    switch (G) {
        case 1:
            err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, k_batch<1, BuiltinModel>, kGroup, 0);
            break;
        case 2:
            err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, k_batch<2, BuiltinModel>, 2 * kGroup, 0);
            break;
        default:
            err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, k_batch<4, BuiltinModel>, 4 * kGroup, 0);
            break;
    }
    MCKL_HIP_TRY(err);
    per_cu = std::max(1, occ);
    if (waves_per_eu > 0)
        per_cu = std::min(per_cu,
            std::max(1, waves_per_eu * 4 / (G * kGroup / 64)));
    return MCKL_HIP_OK;
}

static int set_persistent(mckl_hip_smc_batch *b)
{
    int dev = 0, ncu = 0;
    MCKL_HIP_TRY(hipGetDevice(&dev));
    MCKL_HIP_TRY(hipDeviceGetAttribute(
        &ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t nf = (uint32_t)b->nfilters;
    uint32_t nl = nf;
    if (b->tune.split_tiles > 0) {
        nl = 0;
        while (nl < nf &&
            ceil_div(b->nsorted[nl], kTile) >= (uint64_t)b->tune.split_tiles)
            ++nl;
    }
    b->nlarge = nl;
    int pl = 0, ps = 0;
    int rc = batch_per_cu(std::max(1, b->tune.block / kGroup),
        b->tune.waves_per_eu, pl);
    if (rc == MCKL_HIP_OK)
        rc = batch_per_cu(std::max(1, b->tune.small_block / kGroup),
            b->tune.waves_per_eu, ps);
    if (rc)
        return rc;
    b->persistent = (int)std::min<size_t>((size_t)ncu * pl, nl);
    b->persistent_small = (int)std::min<size_t>((size_t)ncu * ps, nf - nl);
    return MCKL_HIP_OK;
}

extern "C" {

int mckl_hip_smc_batch_create(int model, size_t nfilters, const uint32_t *h_N,
    const double *h_theta, size_t T, const double *h_y, const uint64_t *h_seed,
    double threshold, int scheme, mckl_hip_smc_batch **out)
{
    if (!out || !h_N || !h_theta || !h_y || !h_seed)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    *out = nullptr;
    if (nfilters == 0 || nfilters >= (1u << 31))
        return set_error(MCKL_HIP_EINVAL, "nfilters out of range");
    std::vector<FilterDescT<BuiltinModel>> fd(nfilters);
    std::vector<float> obs(nfilters * T);
    uint64_t total = 0;
    for (size_t f = 0; f < nfilters; ++f) {
        const size_t N = h_N[f];
        int rc = check_common(N, T, threshold, scheme);
        if (rc)
            return rc;
        if (N > (size_t)kMaxTiles * kTile)
            return set_error(MCKL_HIP_EINVAL,
                "batched filters hold at most %d particles", kMaxTiles * kTile);
        ModelConst mc;
        rc = model_const(model, h_theta + 3 * f, mc);
        if (rc)
            return rc;
        SmcArgs &A = fd[f].A;
        A.mc = mc;
        A.k0 = (uint32_t)h_seed[f];
        A.k1 = (uint32_t)(h_seed[f] >> 32);
        A.N = (uint32_t)N;
        A.ntiles = (uint32_t)ceil_div(N, kTile);
        A.tauN = threshold * (double)N;
        A.scheme = scheme;
        A.T = (uint32_t)T;
        A.obs = nullptr; // set after upload
        fd[f].off = total;
        total += (uint64_t)A.ntiles * kTile;
        obs_const(model, h_theta + 3 * f, h_y + T * f, T, obs.data() + T * f);
    }
    // largest filters first (longest-processing-time order)
    std::vector<uint32_t> order(nfilters);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
        [&](uint32_t a, uint32_t b) { return h_N[a] > h_N[b]; });

    mckl_hip_smc_batch *b = new mckl_hip_smc_batch();
    b->nfilters = nfilters;
    b->T = T;
    b->model = model;
    b->y.assign(h_y, h_y + nfilters * T);
    int rc = MCKL_HIP_OK;
#define MCKL_ALLOC(p, n)                                                      \
    do {                                                                      \
        rc = dev_alloc(&(p), (n));                                            \
        if (rc) {                                                             \
            batch_free(b);                                                    \
            return rc;                                                        \
        }                                                                     \
    } while (0)
    MCKL_ALLOC(b->fd, nfilters);
    MCKL_ALLOC(b->order, nfilters);
    MCKL_ALLOC(b->queue, 2);
    MCKL_ALLOC(b->xa, total);
    MCKL_ALLOC(b->xb, total);
    MCKL_ALLOC(b->lw, total);
    MCKL_ALLOC(b->idx, total);
    MCKL_ALLOC(b->cnt, total);
    MCKL_ALLOC(b->E, total);
    MCKL_ALLOC(b->zpos, total);
    if (scheme == MCKL_HIP_MULTINOMIAL || scheme == MCKL_HIP_RESIDUAL)
        MCKL_ALLOC(b->u, total);
    MCKL_ALLOC(b->obs, nfilters * T);
    MCKL_ALLOC(b->logz, nfilters);
#undef MCKL_ALLOC
    for (size_t f = 0; f < nfilters; ++f) {
        fd[f].A.obs = b->obs + T * f;
        fd[f].model = builtin_model(fd[f].A);
    }
    hipError_t err = hipMemcpy(b->fd, fd.data(), nfilters * sizeof(FilterDescT<BuiltinModel>),
        hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(b->order, order.data(), nfilters * sizeof(uint32_t),
            hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(b->obs, obs.data(), obs.size() * sizeof(float),
            hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        rc = hip_status(err, "mckl_hip_smc_batch_create");
        batch_free(b);
        return rc;
    }
    b->nsorted.resize(nfilters);
    for (size_t i = 0; i < nfilters; ++i)
        b->nsorted[i] = h_N[order[i]];
    b->threshold = threshold;
    b->scheme = scheme;
    rc = set_persistent(b);
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipEventCreateWithFlags(&b->done,
            hipEventDisableTiming), "mckl_hip_smc_batch_create");
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipEventCreateWithFlags(&b->fork,
            hipEventDisableTiming), "mckl_hip_smc_batch_create");
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipEventCreateWithFlags(&b->join,
            hipEventDisableTiming), "mckl_hip_smc_batch_create");
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipStreamCreateWithFlags(&b->side,
            hipStreamNonBlocking), "mckl_hip_smc_batch_create");
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipStreamCreateWithFlags(&b->rs,
            hipStreamNonBlocking), "mckl_hip_smc_batch_create");
    if (rc) {
        batch_free(b);
        return rc;
    }
    b->fdh = std::move(fd);
    *out = b;
    return MCKL_HIP_OK;
}

int mckl_hip_smc_batch_set_params(mckl_hip_smc_batch *b,
    const double *h_theta, const uint64_t *h_seed)
{
    if (!b || !h_theta || !h_seed)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    const size_t nf = b->nfilters, T = b->T;
    std::vector<ModelConst> mc(nf);
    for (size_t f = 0; f < nf; ++f) {
        const int rc = model_const(b->model, h_theta + 3 * f, mc[f]);
        if (rc)
            return rc;
    }
    std::vector<float> obs(nf * T);
    for (size_t f = 0; f < nf; ++f) {
        SmcArgs &A = b->fdh[f].A;
        A.mc = mc[f];
        A.k0 = (uint32_t)h_seed[f];
        A.k1 = (uint32_t)(h_seed[f] >> 32);
        b->fdh[f].model = builtin_model(A);
        obs_const(b->model, h_theta + 3 * f, b->y.data() + T * f, T,
            obs.data() + T * f);
    }
    // a run still in flight reads the descriptors: wait for it
    MCKL_HIP_TRY(hipEventSynchronize(b->done));
    MCKL_HIP_TRY(hipMemcpyWithStream(b->fd, b->fdh.data(),
        nf * sizeof(FilterDescT<BuiltinModel>), hipMemcpyHostToDevice, b->rs));
    MCKL_HIP_TRY(hipMemcpyWithStream(b->obs, obs.data(),
        obs.size() * sizeof(float), hipMemcpyHostToDevice, b->rs));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_batch_set_tune(mckl_hip_smc_batch *b, const mckl_hip_tune *t)
{
    if (!b || !t)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (t->block != 256 && t->block != 512 && t->block != 1024)
        return set_error(MCKL_HIP_EINVAL, "block must be 256, 512 or 1024");
    if (t->waves_per_eu < 0 || t->waves_per_eu > 8)
        return set_error(MCKL_HIP_EINVAL, "waves_per_eu must be in 0..8");
    if (t->split_tiles < 0 || t->split_tiles > kMaxTiles)
        return set_error(MCKL_HIP_EINVAL, "split_tiles must be in 0..%d",
            kMaxTiles);
    if (t->split_tiles > 0 && t->small_block != 256 &&
        t->small_block != 512 && t->small_block != 1024)
        return set_error(MCKL_HIP_EINVAL,
            "small_block must be 256, 512 or 1024");
    MCKL_HIP_TRY(hipEventSynchronize(b->done)); // a run in flight uses it
    b->tune = *t;
    if (b->tune.split_tiles == 0)
        b->tune.small_block = 256;
    return set_persistent(b);
}

int mckl_hip_smc_batch_run(mckl_hip_smc_batch *b, int use_graph, void *stream)
{
    if (!b)
        return set_error(MCKL_HIP_EINVAL, "batch is NULL");
    hipStream_t st = as_stream(stream);
    if (!use_graph) {
        const int rc = batch_launch(b, st);
        if (rc)
            return rc;
        MCKL_HIP_TRY(hipEventRecord(b->done, st));
        return MCKL_HIP_OK;
    }
    const long long key =
        (((long long)b->tune.block * 16 + b->tune.waves_per_eu) * 128 +
            b->tune.split_tiles) * 4096 + b->tune.small_block;
    if (!b->exec || b->graph_tune != key) {
        if (b->exec) {
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
        if (b->graph) {
            (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
        }
        hipStream_t cap;
        MCKL_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int rc = MCKL_HIP_OK;
        if (err == hipSuccess) {
            rc = batch_launch(b, cap);
            err = hipStreamEndCapture(cap, &b->graph);
        }
        if (err == hipSuccess && rc == MCKL_HIP_OK)
            err = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
        if (rc)
            return rc;
        MCKL_HIP_TRY(err);
        b->graph_tune = key;
    }
    MCKL_HIP_TRY(hipGraphLaunch(b->exec, st));
    MCKL_HIP_TRY(hipEventRecord(b->done, st));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_batch_result(mckl_hip_smc_batch *b, double *h_loglik)
{
    if (!b || !h_loglik)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    MCKL_HIP_TRY(hipEventSynchronize(b->done));
    MCKL_HIP_TRY(hipMemcpyWithStream(h_loglik, b->logz,
        b->nfilters * sizeof(double), hipMemcpyDeviceToHost, b->rs));
    return MCKL_HIP_OK;
}

} // extern "C"

// ---- autotuner ----
// The launch configurations tried (none changes a bit of the results: every
// filter runs through the same tile stages whatever the workgroup width,
// occupancy or launch it lands in).
static const mckl_hip_tune kTuneCandidates[] = {
    {256, 0, 0, 256}, {256, 2, 0, 256}, {512, 0, 0, 256}, {512, 2, 0, 256},
    {1024, 0, 0, 256}, {512, 0, 2, 256}, {512, 0, 4, 256}, {512, 0, 8, 256},
    {1024, 0, 4, 256}, {1024, 0, 8, 256}, {1024, 0, 16, 256},
    {1024, 0, 24, 256}, {512, 2, 8, 256}, {1024, 2, 8, 256},
};

// Workload signature for the cache: device, model, scheme, T, threshold,
// number of filters and an FNV-1a hash of their sizes in queue order.
static std::string batch_signature(const mckl_hip_smc_batch *b)
{
    int dev = 0;
    hipDeviceProp_t prop;
    std::memset(&prop, 0, sizeof(prop));
    if (hipGetDevice(&dev) != hipSuccess ||
        hipGetDeviceProperties(&prop, dev) != hipSuccess)
        std::strcpy(prop.gcnArchName, "unknown");
    uint64_t h = 1469598103934665603ull;
    for (uint32_t n : b->nsorted) {
        for (int k = 0; k < 4; ++k) {
            h ^= (n >> (8 * k)) & 0xFFu;
            h *= 1099511628211ull;
        }
    }
    char buf[512];
    std::snprintf(buf, sizeof(buf),
        "%s/%dcu/model%d/scheme%d/T%zu/thr%.17g/nf%zu/N%016llx",
        prop.gcnArchName, prop.multiProcessorCount, b->model, b->scheme, b->T,
        b->threshold, b->nfilters, (unsigned long long)h);
    return std::string(buf);
}

static bool json_int(const std::string &line, const char *name, int &v)
{
    const std::string pat = std::string("\"") + name + "\":";
    const size_t p = line.find(pat);
    if (p == std::string::npos)
        return false;
    v = std::atoi(line.c_str() + p + pat.size());
    return true;
}

// the cached configuration of `key` (the last record for it), if any
static bool cache_lookup(const char *path, const std::string &key,
    mckl_hip_tune &t, double &ms)
{
    FILE *f = std::fopen(path, "r");
    if (!f)
        return false;
    bool found = false;
    char buf[2048];
    const std::string kpat = "\"key\": \"" + key + "\"";
    while (std::fgets(buf, sizeof(buf), f)) {
        const std::string line(buf);
        if (line.find(kpat) == std::string::npos)
            continue;
        mckl_hip_tune c{256, 0, 0, 256};
        if (!json_int(line, "block", c.block) ||
            !json_int(line, "waves_per_eu", c.waves_per_eu) ||
            !json_int(line, "split_tiles", c.split_tiles) ||
            !json_int(line, "small_block", c.small_block))
            continue;
        const size_t p = line.find("\"ms\":");
        ms = p == std::string::npos ? 0.0 : std::atof(line.c_str() + p + 5);
        t = c;
        found = true;
    }
    std::fclose(f);
    return found;
}

static void cache_store(const char *path, const std::string &key,
    const mckl_hip_tune &t, double ms)
{
    FILE *f = std::fopen(path, "a");
    if (!f)
        return;
    std::fprintf(f,
        "{\"key\": \"%s\", \"block\": %d, \"waves_per_eu\": %d, "
        "\"split_tiles\": %d, \"small_block\": %d, \"ms\": %.4f}\n",
        key.c_str(), t.block, t.waves_per_eu, t.split_tiles, t.small_block,
        ms);
    std::fclose(f);
}

// one graph run timed with events on st (after a warm-up run that also
// captures the configuration's graph); the filters' log-likelihoods to h_ll
static int time_run(mckl_hip_smc_batch *b, hipStream_t st, double *h_ll,
    double &ms)
{
    int rc = mckl_hip_smc_batch_run(b, 1, st);
    if (rc)
        return rc;
    hipEvent_t e0, e1;
    MCKL_HIP_TRY(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_status(err, "hipEventCreate");
    }
    err = hipEventRecord(e0, st);
    if (err == hipSuccess) {
        rc = mckl_hip_smc_batch_run(b, 1, st);
        if (rc == MCKL_HIP_OK)
            err = hipEventRecord(e1, st);
    }
    if (rc == MCKL_HIP_OK && err == hipSuccess)
        err = hipEventSynchronize(e1);
    float f = 0.0f;
    if (rc == MCKL_HIP_OK && err == hipSuccess)
        err = hipEventElapsedTime(&f, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc)
        return rc;
    MCKL_HIP_TRY(err);
    ms = f;
    return mckl_hip_smc_batch_result(b, h_ll);
}

extern "C" {

int mckl_hip_smc_batch_autotune(mckl_hip_smc_batch *b, const char *cache_path,
    mckl_hip_tune *best, double *best_ms, void *stream)
{
    if (!b)
        return set_error(MCKL_HIP_EINVAL, "batch is NULL");
    hipStream_t st = as_stream(stream);
    const std::string key = batch_signature(b);
    mckl_hip_tune t{256, 0, 0, 256};
    double ms = 0.0;
    if (cache_path && cache_lookup(cache_path, key, t, ms)) {
        const int rc = mckl_hip_smc_batch_set_tune(b, &t);
        if (rc)
            return rc;
    } else {
        std::vector<double> ref(b->nfilters), got(b->nfilters);
        bool have = false;
        for (const mckl_hip_tune &c : kTuneCandidates) {
            int rc = mckl_hip_smc_batch_set_tune(b, &c);
            double cms = 0.0;
            if (rc == MCKL_HIP_OK)
                rc = time_run(b, st, have ? got.data() : ref.data(), cms);
            if (rc)
                return rc;
            if (have && std::memcmp(got.data(), ref.data(),
                            b->nfilters * sizeof(double)) != 0)
                return set_error(MCKL_HIP_EHIP,
                    "autotune: configuration (%d, %d, %d, %d) changed the "
                    "results", c.block, c.waves_per_eu, c.split_tiles,
                    c.small_block);
            if (!have || cms < ms) {
                t = c;
                ms = cms;
            }
            have = true;
        }
        const int rc = mckl_hip_smc_batch_set_tune(b, &t);
        if (rc)
            return rc;
        if (cache_path)
            cache_store(cache_path, key, t, ms);
    }
    if (best)
        *best = t;
    if (best_ms)
        *best_ms = ms;
    return MCKL_HIP_OK;
}

int mckl_hip_smc_batch_destroy(mckl_hip_smc_batch *b)
{
    if (b)
        batch_free(b);
    return MCKL_HIP_OK;
}

} // extern "C"
