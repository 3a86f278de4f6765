// Models written as a caller of mckl_amd/include/mckl_hip_fused.hpp would
// write them, run through the fused plain and batched samplers, behind a C
// interface for tests/test_gpu_fused.py:
//  * UserLG: the linear-Gaussian model's arithmetic, written out here (it
//    must give the built-in sampler's bits);
//  * Growth: the nonstationary growth model, x_t = (0.5 x + 25 x / (1 +
//    x^2)) + u_t + s e, y_t = x^2 / 20 + sy w (the oracle's model 3);
//  * CV2: a two-component state (position, velocity), the oracle's model 4.
// Test infrastructure only.
#include "mckl_hip_fused.hpp"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

namespace {

struct UserLG {
    float a, s, sd0, inv_sy, cst;
    const float *obs;
    struct Step {
        float a, s, inv_sy, cst, y;
        __device__ float move(float x, float e) const
        {
            return __fadd_rn(__fmul_rn(a, x), __fmul_rn(s, e));
        }
        __device__ float log_weight(float x) const
        {
            const float z = __fmul_rn(__fsub_rn(y, x), inv_sy);
            return __fadd_rn(__fmul_rn(-0.5f, __fmul_rn(z, z)), cst);
        }
    };
    __device__ float init(float e) const { return __fmul_rn(sd0, e); }
    __device__ Step at(uint32_t t) const
    {
        return Step{a, s, inv_sy, cst, obs[t - 1]};
    }
};

struct Growth {
    float s, sd0, inv_sy, cst;
    const float *obs, *drift;
    struct Step {
        float s, inv_sy, cst, y, u;
        __device__ float move(float x, float e) const
        {
            const float q = __fdiv_rn(x, __fadd_rn(1.0f, __fmul_rn(x, x)));
            const float mean = __fadd_rn(
                __fadd_rn(__fmul_rn(0.5f, x), __fmul_rn(25.0f, q)), u);
            return __fadd_rn(mean, __fmul_rn(s, e));
        }
        __device__ float log_weight(float x) const
        {
            const float h = __fmul_rn(__fmul_rn(x, x), 0.05f);
            const float z = __fmul_rn(__fsub_rn(y, h), inv_sy);
            return __fadd_rn(__fmul_rn(-0.5f, __fmul_rn(z, z)), cst);
        }
    };
    __device__ float init(float e) const { return __fmul_rn(sd0, e); }
    __device__ Step at(uint32_t t) const
    {
        return Step{s, inv_sy, cst, obs[t - 1], drift[t - 1]};
    }
};

// a vector state: (position, velocity), position += velocity + sp e0,
// velocity += sv e1, y = position + sy w (the oracle's model 4)
struct CV2 {
    static constexpr int dim = 2;
    float sp, sv, inv_sy, cst;
    const float *obs;
    struct Step {
        float sp, sv, inv_sy, cst, y;
        __device__ void move(const float *xp, const float *e, float *xn) const
        {
            xn[0] = __fadd_rn(__fadd_rn(xp[0], xp[1]), __fmul_rn(sp, e[0]));
            xn[1] = __fadd_rn(xp[1], __fmul_rn(sv, e[1]));
        }
        __device__ float log_weight(const float *x) const
        {
            const float z = __fmul_rn(__fsub_rn(y, x[0]), inv_sy);
            return __fadd_rn(__fmul_rn(-0.5f, __fmul_rn(z, z)), cst);
        }
    };
    __device__ void init(const float *e, float *x) const
    {
        x[0] = e[0];
        x[1] = e[1];
    }
    __device__ Step at(uint32_t t) const
    {
        return Step{sp, sv, inv_sy, cst, obs[t - 1]};
    }
};

// The reference's particle-filter example (example/algorithm/include/
// algorithm_pf.hpp:91-150) as a four-float state (p, q, s, t): iteration 0
// draws p, q ~ N(0, 2), s, t ~ N(0, 1) and weighs the first observation;
// every later iteration moves p += 0.1 s + sqrt(0.02) e, s += sqrt(0.001) e
// (then q, t likewise) and weighs by -5.5 (log1p(0.1 (10 (p - x))^2) +
// log1p(0.1 (10 (q - y))^2)).  The fused samplers move before every
// weighting, so step 1's move is the identity.  FP32 states and the
// samplers' own noise stream: the same model, not the reference's bits.
struct PF4 {
    static constexpr int dim = 4;
    const float *ox, *oy;
    struct Step {
        float ox, oy;
        bool first;
        __device__ void move(const float *xp, const float *e, float *xn) const
        {
            if (first) {
                for (int d = 0; d < 4; ++d)
                    xn[d] = xp[d];
                return;
            }
            const float sp = 0.141421356f, sv = 0.0316227766f;
            xn[0] = xp[0] + (0.1f * xp[2] + sp * e[0]);
            xn[2] = xp[2] + sv * e[1];
            xn[1] = xp[1] + (0.1f * xp[3] + sp * e[2]);
            xn[3] = xp[3] + sv * e[3];
        }
        __device__ float log_weight(const float *x) const
        {
            const float a = 10.0f * (x[0] - ox), b = 10.0f * (x[1] - oy);
            return -5.5f * (log1pf(0.1f * a * a) + log1pf(0.1f * b * b));
        }
    };
    __device__ void init(const float *e, float *x) const
    {
        x[0] = 2.0f * e[0];
        x[1] = 2.0f * e[1];
        x[2] = e[2];
        x[3] = e[3];
    }
    __device__ Step at(uint32_t t) const
    {
        return Step{ox[t - 1], oy[t - 1], t == 1};
    }
};

const double kL2pi = std::log(2.0 * 3.141592653589793);

UserLG make_lg(const double *th, const float *d_obs)
{
    return UserLG{(float)th[0], (float)th[1], 1.0f, (float)(1.0 / th[2]),
        (float)(-std::log(th[2]) - 0.5 * kL2pi), d_obs};
}

// theta = (s, sy, sd0)
Growth make_growth(const double *th, const float *d_obs, const float *d_u)
{
    return Growth{(float)th[0], (float)th[2], (float)(1.0 / th[1]),
        (float)(-std::log(th[1]) - 0.5 * kL2pi), d_obs, d_u};
}

// theta = (sp, sv, sy)
CV2 make_cv2(const double *th, const float *d_obs)
{
    return CV2{(float)th[0], (float)th[1], (float)(1.0 / th[2]),
        (float)(-std::log(th[2]) - 0.5 * kL2pi), d_obs};
}

std::string g_msg;

// per filter: T observations (model 1) or T observations and T drifts
// (model 3) from y, as floats on the device
struct Obs {
    std::vector<float *> p;
    ~Obs()
    {
        for (float *q : p)
            (void)hipFree(q);
    }
    const float *put(const double *y, size_t n)
    {
        std::vector<float> h(n);
        for (size_t i = 0; i < n; ++i)
            h[i] = (float)y[i];
        float *d = nullptr;
        if (hipMalloc(&d, n * sizeof(float)) != hipSuccess)
            throw mckl::RuntimeAssert("hipMalloc");
        p.push_back(d);
        if (hipMemcpy(d, h.data(), n * sizeof(float), hipMemcpyHostToDevice) !=
            hipSuccess)
            throw mckl::RuntimeAssert("hipMemcpy");
        return d;
    }
};

template <typename M>
void plain(const M &model, size_t T, size_t N, uint64_t seed,
    double threshold, int scheme, double *ll, double *ess, int32_t *res)
{
    mckl::FusedSMCSampler<M> s(model, N, T, seed, threshold, scheme);
    s.iterate();
    s.iterate(); // a replay of the captured graph gives the same run
    *ll = s.log_likelihood();
    const std::vector<double> e = s.ess_history();
    const std::vector<int32_t> r = s.resampled_history();
    for (size_t t = 0; t < T; ++t) {
        ess[t] = e[t];
        res[t] = r[t];
    }
}

template <typename M>
void plain_moments(const M &model, size_t T, size_t N, uint64_t seed,
    double threshold, int scheme, int k, double *ll, double *est)
{
    mckl::FusedSMCSampler<M> s(model, N, T, seed, threshold, scheme);
    s.estimate_moments(k);
    s.iterate();
    *ll = s.log_likelihood();
    const std::vector<double> e = s.estimates();
    for (size_t i = 0; i < e.size(); ++i)
        est[i] = e[i];
}

} // namespace

extern "C" {

const char *fused_last_error() { return g_msg.c_str(); }

// model 1 (UserLG), 3 (Growth) or 4 (CV2); y: T (models 1, 4) or 2 T
// (model 3) values
int fused_plain_run(int model, const double *theta, const double *y,
    size_t T, size_t N, uint64_t seed, double threshold, int scheme,
    double *ll, double *ess, int32_t *res)
{
    try {
        Obs o;
        if (model == 1)
            plain(make_lg(theta, o.put(y, T)), T, N, seed, threshold, scheme,
                ll, ess, res);
        else if (model == 4)
            plain(make_cv2(theta, o.put(y, T)), T, N, seed, threshold, scheme,
                ll, ess, res);
        else
            plain(make_growth(theta, o.put(y, T), o.put(y + T, T)), T, N,
                seed, threshold, scheme, ll, ess, res);
        return 0;
    } catch (const std::exception &e) {
        g_msg = e.what();
        return -1;
    }
}

// the run's selection estimates of x^1..x^k, T x k step-major
int fused_plain_moments(int model, const double *theta, const double *y,
    size_t T, size_t N, uint64_t seed, double threshold, int scheme, int k,
    double *ll, double *est)
{
    try {
        Obs o;
        if (model == 1)
            plain_moments(make_lg(theta, o.put(y, T)), T, N, seed, threshold,
                scheme, k, ll, est);
        else if (model == 4)
            plain_moments(make_cv2(theta, o.put(y, T)), T, N, seed,
                threshold, scheme, k, ll, est);
        else
            plain_moments(make_growth(theta, o.put(y, T), o.put(y + T, T)),
                T, N, seed, threshold, scheme, k, ll, est);
        return 0;
    } catch (const std::exception &e) {
        g_msg = e.what();
        return -1;
    }
}

// the particle-filter example on the fused plain sampler (stratified
// resampling below ESS N/2, as the example): xy holds the T observations as
// (x, y) pairs; writes log Z-hat, the ESS and resampling histories, the
// last step's posterior mean and second moment of p (post[0], post[1]) and
// the time of one iterate() in ms
int fused_pf_run(const double *xy, size_t T, size_t N, uint64_t seed,
    double *ll, double *ess, int32_t *res, double *post, double *ms)
{
    try {
        Obs o;
        std::vector<double> xs(T), ys(T);
        for (size_t t = 0; t < T; ++t) {
            xs[t] = xy[2 * t];
            ys[t] = xy[2 * t + 1];
        }
        const PF4 model{o.put(xs.data(), T), o.put(ys.data(), T)};
        mckl::FusedSMCSampler<PF4> s(model, N, T, seed, 0.5,
            MCKL_HIP_STRATIFIED);
        s.iterate();
        (void)s.log_likelihood();
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) !=
            hipSuccess)
            throw mckl::RuntimeAssert("hipEventCreate");
        (void)hipEventRecord(a, nullptr);
        s.iterate();
        (void)hipEventRecord(b, nullptr);
        (void)hipEventSynchronize(b);
        float f = 0.0f;
        (void)hipEventElapsedTime(&f, a, b);
        *ms = f;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *ll = s.log_likelihood();
        const std::vector<double> e = s.ess_history();
        const std::vector<int32_t> r = s.resampled_history();
        for (size_t t = 0; t < T; ++t) {
            ess[t] = e[t];
            res[t] = r[t];
        }
        // moments of the first component (p) at every step
        s.estimate_moments(2);
        s.iterate();
        const std::vector<double> m = s.estimates();
        post[0] = m[2 * (T - 1)];
        post[1] = m[2 * (T - 1) + 1];
        return 0;
    } catch (const std::exception &e) {
        g_msg = e.what();
        return -1;
    }
}

// nf filters with their own theta (3 each), y (T or 2 T each), N and seed
// (models 1, 3, 4)
int fused_batch_run(int model, size_t nf, const uint32_t *N,
    const double *theta, const double *y, size_t T, const uint64_t *seeds,
    double threshold, int scheme, double *ll)
{
    try {
        Obs o;
        const std::vector<uint32_t> n(N, N + nf);
        const std::vector<uint64_t> sd(seeds, seeds + nf);
        std::vector<double> out;
        if (model == 1) {
            std::vector<UserLG> m;
            for (size_t f = 0; f < nf; ++f)
                m.push_back(make_lg(theta + 3 * f, o.put(y + T * f, T)));
            mckl::FusedSMCBatch<UserLG> b(m, n, T, sd, threshold, scheme);
            b.iterate();
            out = b.log_likelihood();
        } else if (model == 4) {
            std::vector<CV2> m;
            for (size_t f = 0; f < nf; ++f)
                m.push_back(make_cv2(theta + 3 * f, o.put(y + T * f, T)));
            mckl::FusedSMCBatch<CV2> b(m, n, T, sd, threshold, scheme);
            b.iterate();
            b.iterate();
            out = b.log_likelihood();
        } else {
            std::vector<Growth> m;
            for (size_t f = 0; f < nf; ++f)
                m.push_back(make_growth(theta + 3 * f,
                    o.put(y + 2 * T * f, T), o.put(y + 2 * T * f + T, T)));
            mckl::FusedSMCBatch<Growth> b(m, n, T, sd, threshold, scheme);
            b.iterate();
            b.iterate();
            out = b.log_likelihood();
        }
        for (size_t f = 0; f < nf; ++f)
            ll[f] = out[f];
        return 0;
    } catch (const std::exception &e) {
        g_msg = e.what();
        return -1;
    }
}

} // extern "C"
