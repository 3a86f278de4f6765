// mckl_hip.hpp -- C++ host API of the MI355X hot path, mirroring the
// reference's interface (namespace mckl, same type and function names) over
// the C ABI in include/mckl_hip.h.  Header-only; link libmckl_hip.so.
//
// Differences from the reference, all forced by the device boundary:
//  * output and weight arrays are device pointers;
//  * the scalar engine call operator()() and ess() synchronise (the value is
//    produced on the device);
//  * errors raise mckl::RuntimeAssert carrying the C ABI's message, as the
//    reference does with MCKL_RUNTIME_ASSERT_AS_EXCEPTION
//    (internal/assert.hpp:49-84).
#ifndef MCKL_HIP_HPP
#define MCKL_HIP_HPP

#include <mckl_hip.h>
#include <hip/hip_runtime_api.h>

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mckl {

class RuntimeAssert : public std::runtime_error {
  public:
    explicit RuntimeAssert(const std::string &msg) : std::runtime_error(msg) {}
};

inline void hip_check(int rc)
{
    if (rc != MCKL_HIP_OK)
        throw RuntimeAssert(std::string("mckl_hip: ") + mckl_hip_last_error());
}

// stream used by all calls of this thread (hipStream_t; nullptr = default)
inline void *&current_stream()
{
    static thread_local void *s = nullptr;
    return s;
}

/// CounterEngine<ResultType, Generator> (random/counter.hpp:61-288)
template <int Kind, typename ResultType>
class CounterEngineHIP
{
  public:
    using result_type = ResultType;
    using skip_type = std::uint64_t;
    using key_type = typename std::conditional<Kind == MCKL_HIP_THREEFRY4X64,
        std::array<std::uint64_t, 4>,
        typename std::conditional<Kind == MCKL_HIP_PHILOX4X32,
            std::array<std::uint32_t, 2>,
            std::array<std::uint32_t, 4>>::type>::type;
    using ctr_type = typename std::conditional<Kind == MCKL_HIP_THREEFRY4X64,
        std::array<std::uint64_t, 4>, std::array<std::uint64_t, 2>>::type;

    explicit CounterEngineHIP(result_type s = 1)
    {
        key_type k{};
        reset(k);
        seed(s);
    }

    explicit CounterEngineHIP(const key_type &k) { reset(k); }

    void seed(result_type s) { hip_check(mckl_hip_engine_seed(&state_, s)); }

    void seed(const key_type &k) { reset(k); }

    key_type key() const
    {
        key_type k;
        for (std::size_t i = 0; i != k.size(); ++i)
            k[i] = static_cast<typename key_type::value_type>(state_.key[i]);
        return k;
    }

    ctr_type ctr() const
    {
        ctr_type c;
        for (std::size_t i = 0; i != c.size(); ++i)
            c[i] = state_.ctr[i];
        return c;
    }

    void ctr(const ctr_type &c)
    {
        std::uint64_t w[4] = {0, 0, 0, 0};
        for (std::size_t i = 0; i != c.size(); ++i)
            w[i] = c[i];
        hip_check(mckl_hip_engine_set_ctr(&state_, w));
    }

    result_type operator()()
    {
        std::uint64_t v = 0;
        hip_check(mckl_hip_engine_next(&state_, &v));
        return static_cast<result_type>(v);
    }

    void discard(skip_type n) { hip_check(mckl_hip_engine_discard(&state_, n)); }

    static constexpr result_type min() { return 0; }
    static constexpr result_type max() { return ~result_type(0); }

    mckl_hip_engine &state() { return state_; }
    const mckl_hip_engine &state() const { return state_; }

    friend bool operator==(const CounterEngineHIP &a, const CounterEngineHIP &b)
    {
        return a.state_.index == b.state_.index && a.ctr() == b.ctr() &&
            a.key() == b.key();
    }
    friend bool operator!=(const CounterEngineHIP &a, const CounterEngineHIP &b)
    {
        return !(a == b);
    }

    /// operator<< (random/counter.hpp:235-249): result_ ctr_ generator_
    /// index_, arrays as "size v0 v1 ..."; result_ is the buffered block
    /// G(ctr), computed on the device
    template <typename CharT, typename Traits>
    friend std::basic_ostream<CharT, Traits> &operator<<(
        std::basic_ostream<CharT, Traits> &os, const CounterEngineHIP &eng)
    {
        if (!os)
            return os;
        constexpr std::size_t nbytes = Kind == MCKL_HIP_THREEFRY4X64 ? 32 : 16;
        unsigned char blk[nbytes];
        hip_check(mckl_hip_engine_block(&eng.state_, blk));
        constexpr std::size_t M = nbytes / sizeof(ResultType);
        os << M;
        for (std::size_t i = 0; i != M; ++i) {
            ResultType v;
            std::memcpy(&v, blk + i * sizeof(ResultType), sizeof(v));
            os << ' ' << +v;
        }
        const ctr_type c = eng.ctr();
        os << ' ' << c.size();
        for (auto v : c)
            os << ' ' << v;
        const key_type k = eng.key();
        if (Kind == MCKL_HIP_THREEFRY4X64) {
            std::uint64_t par = 0x1BD11BDAA9FC1A22ULL;
            for (auto v : k)
                par ^= v;
            os << ' ' << 8;
            for (auto v : k)
                os << ' ' << v;
            os << ' ' << par << " 0 0 0";
        } else {
            os << ' ' << k.size();
            for (auto v : k)
                os << ' ' << v;
        }
        os << ' ' << eng.state_.index;
        return os;
    }

    /// operator>> (random/counter.hpp:251-271)
    template <typename CharT, typename Traits>
    friend std::basic_istream<CharT, Traits> &operator>>(
        std::basic_istream<CharT, Traits> &is, CounterEngineHIP &eng)
    {
        if (!is)
            return is;
        std::size_t m = 0, kc = 0, kk = 0;
        unsigned index = 0;
        is >> std::ws >> m;
        for (std::size_t i = 0; i != m && is; ++i) {
            unsigned long long r;
            is >> std::ws >> r;
        }
        ctr_type c{};
        is >> std::ws >> kc;
        for (std::size_t i = 0; i != kc && is; ++i) {
            unsigned long long v;
            is >> std::ws >> v;
            if (i < c.size())
                c[i] = v;
        }
        key_type k{};
        is >> std::ws >> kk;
        for (std::size_t i = 0; i != kk && is; ++i) {
            unsigned long long v;
            is >> std::ws >> v;
            if (i < k.size())
                k[i] = static_cast<typename key_type::value_type>(v);
        }
        is >> std::ws >> index;
        if (is && kc == c.size() && index <= m) {
            CounterEngineHIP tmp(k);
            tmp.ctr(c);
            tmp.state_.index = index;
            eng = tmp;
        } else {
            is.setstate(std::ios_base::failbit);
        }
        return is;
    }

  private:
    mckl_hip_engine state_;

    void reset(const key_type &k)
    {
        std::uint64_t w[4] = {0, 0, 0, 0};
        for (std::size_t i = 0; i != k.size(); ++i)
            w[i] = k[i];
        hip_check(mckl_hip_engine_init(
            &state_, Kind, int(sizeof(ResultType) * 8), w));
    }
};

using Philox4x32 = CounterEngineHIP<MCKL_HIP_PHILOX4X32, std::uint32_t>;
using Philox4x32_64 = CounterEngineHIP<MCKL_HIP_PHILOX4X32, std::uint64_t>;
using Threefry4x64 = CounterEngineHIP<MCKL_HIP_THREEFRY4X64, std::uint32_t>;
using Threefry4x64_64 = CounterEngineHIP<MCKL_HIP_THREEFRY4X64, std::uint64_t>;
using ARS = CounterEngineHIP<MCKL_HIP_ARS, std::uint32_t>;
using ARS_64 = CounterEngineHIP<MCKL_HIP_ARS, std::uint64_t>;

/// rand(rng, n, r): random/counter.hpp:298-303 (r in device memory)
template <int K, typename R>
inline void rand(CounterEngineHIP<K, R> &rng, std::size_t n, R *d_r)
{
    hip_check(mckl_hip_rand(&rng.state(), n, d_r, current_stream()));
}

template <typename RealType, int Interval>
class U01DistributionHIP
{
  public:
    using result_type = RealType;
    template <int K, typename R>
    void operator()(CounterEngineHIP<K, R> &rng, std::size_t n, RealType *d_r)
    {
        fill(rng.state(), n, d_r);
    }

  private:
    static void fill(mckl_hip_engine &e, std::size_t n, double *r)
    {
        hip_check(mckl_hip_rand_u01_f64(&e, Interval, n, r, current_stream()));
    }
    static void fill(mckl_hip_engine &e, std::size_t n, float *r)
    {
        hip_check(mckl_hip_rand_u01_f32(&e, Interval, n, r, current_stream()));
    }
};

template <typename T>
using U01CCDistribution = U01DistributionHIP<T, MCKL_HIP_U01_CC>;
template <typename T>
using U01CODistribution = U01DistributionHIP<T, MCKL_HIP_U01_CO>;
template <typename T>
using U01OCDistribution = U01DistributionHIP<T, MCKL_HIP_U01_OC>;
template <typename T>
using U01OODistribution = U01DistributionHIP<T, MCKL_HIP_U01_OO>;
template <typename T>
using U01CanonicalDistribution = U01DistributionHIP<T, MCKL_HIP_U01_CANONICAL>;

/// NormalDistribution<RealType> (random/normal_distribution.hpp:113-154)
template <typename RealType>
class NormalDistribution
{
  public:
    using result_type = RealType;
    explicit NormalDistribution(RealType mean = 0, RealType stddev = 1)
        : mean_(mean), stddev_(stddev)
    {
    }
    RealType mean() const { return mean_; }
    RealType stddev() const { return stddev_; }

    template <int K, typename R>
    void operator()(CounterEngineHIP<K, R> &rng, std::size_t n, RealType *d_r)
    {
        fill(rng.state(), n, d_r);
    }

  private:
    RealType mean_, stddev_;
    void fill(mckl_hip_engine &e, std::size_t n, float *r)
    {
        hip_check(mckl_hip_rand_normal_f32(
            &e, n, mean_, stddev_, r, current_stream()));
    }
    void fill(mckl_hip_engine &e, std::size_t n, double *r)
    {
        hip_check(mckl_hip_rand_normal_f64(
            &e, n, mean_, stddev_, r, current_stream()));
    }
};

/// UniformRealDistribution<RealType> (random/uniform_real_distribution.hpp:
/// 65-95); a < b, else RuntimeAssert at construction
template <typename RealType>
class UniformRealDistribution
{
  public:
    using result_type = RealType;
    explicit UniformRealDistribution(RealType a = 0, RealType b = 1)
        : a_(a), b_(b)
    {
        if (!(a < b))
            throw RuntimeAssert("**UniformRealDistribution** constructed with "
                                "invalid arguments");
    }
    RealType a() const { return a_; }
    RealType b() const { return b_; }
    RealType min() const { return a_; }
    RealType max() const { return b_; }

    template <int K, typename R>
    void operator()(CounterEngineHIP<K, R> &rng, std::size_t n, RealType *d_r)
    {
        fill(rng.state(), n, d_r);
    }

  private:
    RealType a_, b_;
    void fill(mckl_hip_engine &e, std::size_t n, float *r)
    {
        hip_check(mckl_hip_rand_uniform_real_f32(
            &e, n, a_, b_, r, current_stream()));
    }
    void fill(mckl_hip_engine &e, std::size_t n, double *r)
    {
        hip_check(mckl_hip_rand_uniform_real_f64(
            &e, n, a_, b_, r, current_stream()));
    }
};

/// rand(rng, dist, n, r): random/internal/common.hpp:706-712
template <int K, typename R, typename Dist>
inline void rand(CounterEngineHIP<K, R> &rng, Dist &dist, std::size_t n,
    typename Dist::result_type *d_r)
{
    dist(rng, n, d_r);
}

/// Weight on device memory (core/weight.hpp:44-228).  The caller owns the
/// device array of n doubles and a device array of 2 doubles for
/// {ESS, log normaliser increment}.
class Weight
{
  public:
    Weight(std::size_t n, double *d_w, double *d_stats)
        : n_(n), w_(d_w), stats_(d_stats)
    {
        set_equal();
    }
    std::size_t size() const { return n_; }
    double *data() { return w_; }
    void set_equal()
    {
        hip_check(mckl_hip_weight_set_equal(n_, w_, stats_, current_stream()));
    }
    void set_log(const double *d_v)
    {
        hip_check(mckl_hip_weight_set_log(n_, d_v, w_, stats_, current_stream()));
    }
    void add_log(const double *d_v)
    {
        hip_check(mckl_hip_weight_add_log(n_, d_v, w_, stats_, current_stream()));
    }
    const double *stats() const { return stats_; }
    /// ESS of the last set/add (weight.hpp:64); synchronises the stream
    double ess() const
    {
        double s[2];
        if (hipStreamSynchronize(static_cast<hipStream_t>(current_stream())) !=
                hipSuccess ||
            hipMemcpy(s, stats_, sizeof(s), hipMemcpyDeviceToHost) !=
                hipSuccess)
            throw RuntimeAssert("mckl_hip: reading the ESS failed");
        return s[0];
    }

  private:
    std::size_t n_;
    double *w_;
    double *stats_;
};

enum ResampleScheme {
    Multinomial = MCKL_HIP_MULTINOMIAL,
    Stratified = MCKL_HIP_STRATIFIED,
    Systematic = MCKL_HIP_SYSTEMATIC,
    Residual = MCKL_HIP_RESIDUAL,
    ResidualStratified = MCKL_HIP_RESIDUAL_STRATIFIED,
    ResidualSystematic = MCKL_HIP_RESIDUAL_SYSTEMATIC
};

/// ResampleAlgorithm (algorithm/resample.hpp:515-576) on device arrays
template <ResampleScheme Scheme>
class ResampleAlgorithmHIP
{
  public:
    template <int K, typename R>
    void operator()(std::size_t N, std::size_t M, CounterEngineHIP<K, R> &rng,
        const double *d_w, std::uint32_t *d_rep,
        std::uint32_t *d_idx = nullptr) const
    {
        hip_check(mckl_hip_resample(Scheme, N, M, &rng.state(), d_w, d_rep,
            d_idx, current_stream()));
    }
};

using ResampleMultinomial = ResampleAlgorithmHIP<Multinomial>;
using ResampleStratified = ResampleAlgorithmHIP<Stratified>;
using ResampleResidual = ResampleAlgorithmHIP<Residual>;
using ResampleResidualStratified = ResampleAlgorithmHIP<ResidualStratified>;
using ResampleResidualSystematic = ResampleAlgorithmHIP<ResidualSystematic>;
using ResampleSystematic = ResampleAlgorithmHIP<Systematic>;

inline void resample_trans_u01_rep(std::size_t N, std::size_t M,
    const double *d_w, const double *d_u01, std::uint32_t *d_rep)
{
    hip_check(mckl_hip_resample_trans_u01_rep(
        N, M, d_w, d_u01, d_rep, current_stream()));
}

inline void resample_trans_rep_index(std::size_t N, std::size_t M,
    const std::uint32_t *d_rep, std::uint32_t *d_idx)
{
    hip_check(mckl_hip_resample_trans_rep_index(
        N, M, d_rep, d_idx, current_stream()));
}

enum class MatrixLayout { RowMajor = MCKL_HIP_ROW_MAJOR,
    ColMajor = MCKL_HIP_COL_MAJOR };
constexpr MatrixLayout RowMajor = MatrixLayout::RowMajor;
constexpr MatrixLayout ColMajor = MatrixLayout::ColMajor;

/// ResampleIndex<std::uint32_t> (algorithm/resample.hpp:597-832); index
/// vectors and matrices are device arrays, index(id, ...) returns to the host
class ResampleIndexHIP
{
  public:
    using index_type = std::uint32_t;

    ResampleIndexHIP() { hip_check(mckl_hip_resample_index_create(&h_)); }
    ~ResampleIndexHIP() { mckl_hip_resample_index_destroy(h_); }
    ResampleIndexHIP(const ResampleIndexHIP &) = delete;
    ResampleIndexHIP &operator=(const ResampleIndexHIP &) = delete;

    std::size_t num_iter() const
    {
        std::size_t n = 0;
        hip_check(mckl_hip_resample_index_num_iter(h_, &n));
        return n;
    }

    std::size_t size() const { return size(SIZE_MAX); }

    std::size_t size(std::size_t iter) const
    {
        std::size_t n = 0;
        hip_check(mckl_hip_resample_index_size(h_, iter, &n));
        return n;
    }

    void reset() { hip_check(mckl_hip_resample_index_reset(h_)); }
    void clear() { hip_check(mckl_hip_resample_index_clear(h_)); }

    void push_back(std::size_t N)
    {
        hip_check(mckl_hip_resample_index_push_back(
            h_, N, nullptr, current_stream()));
    }

    void push_back(std::size_t N, const index_type *d_first)
    {
        hip_check(mckl_hip_resample_index_push_back(
            h_, N, d_first, current_stream()));
    }

    index_type index(std::size_t id) const
    {
        return index(id, num_iter() - 1, 0);
    }

    index_type index(std::size_t id, std::size_t iter_back) const
    {
        return index(id, iter_back, 0);
    }

    index_type index(
        std::size_t id, std::size_t iter_back, std::size_t iter) const
    {
        index_type *d = nullptr;
        index_type r = static_cast<index_type>(id);
        if (hipMalloc(reinterpret_cast<void **>(&d), 2 * sizeof(r)) !=
            hipSuccess)
            throw RuntimeAssert("mckl_hip: hipMalloc failed");
        int rc = hipMemcpy(d, &r, sizeof(r), hipMemcpyHostToDevice) ==
                hipSuccess ?
            mckl_hip_resample_index_index(
                h_, 1, d, iter_back, iter, d + 1, current_stream()) :
            MCKL_HIP_EHIP;
        if (rc == MCKL_HIP_OK &&
            hipMemcpy(&r, d + 1, sizeof(r), hipMemcpyDeviceToHost) !=
                hipSuccess)
            rc = MCKL_HIP_EHIP;
        hipFree(d);
        hip_check(rc);
        return r;
    }

    std::size_t index_matrix_nrow() const
    {
        return index_matrix_nrow(num_iter() - 1);
    }

    std::size_t index_matrix_nrow(std::size_t iter_back) const
    {
        std::size_t r = 0, c = 0;
        hip_check(mckl_hip_resample_index_matrix_shape(h_, iter_back, 0, &r, &c));
        return r;
    }

    std::size_t index_matrix_ncol() const
    {
        return index_matrix_ncol(num_iter() - 1, 0);
    }

    std::size_t index_matrix_ncol(std::size_t iter_back) const
    {
        return index_matrix_ncol(iter_back, 0);
    }

    std::size_t index_matrix_ncol(
        std::size_t iter_back, std::size_t iter) const
    {
        std::size_t r = 0, c = 0;
        hip_check(
            mckl_hip_resample_index_matrix_shape(h_, iter_back, iter, &r, &c));
        return c;
    }

    /// writes nrow * ncol entries to d_first, returns d_first + nrow * ncol
    index_type *read_index_matrix(MatrixLayout layout, index_type *d_first,
        std::size_t iter_back, std::size_t iter) const
    {
        hip_check(mckl_hip_resample_index_read_matrix(h_,
            static_cast<int>(layout), iter_back, iter, d_first,
            current_stream()));
        return d_first +
            index_matrix_nrow(iter_back) * index_matrix_ncol(iter_back, iter);
    }

    index_type *read_index_matrix(MatrixLayout layout, index_type *d_first,
        std::size_t iter_back) const
    {
        return read_index_matrix(layout, d_first, iter_back, 0);
    }

    index_type *read_index_matrix(
        MatrixLayout layout, index_type *d_first) const
    {
        return read_index_matrix(layout, d_first, num_iter() - 1, 0);
    }

  private:
    mckl_hip_resample_index *h_ = nullptr;
};

using ResampleIndex = ResampleIndexHIP;

/// SMCSampler (algorithm/smc.hpp:151-421) for the built-in models
class SMCSamplerHIP
{
  public:
    SMCSamplerHIP(int model, const std::array<double, 3> &theta,
        const std::vector<double> &y, std::size_t N, std::uint64_t seed = 1)
        : model_(model), theta_(theta), y_(y), N_(N), seed_(seed)
    {
    }
    ~SMCSamplerHIP()
    {
        if (s_)
            mckl_hip_smc_destroy(s_);
    }
    SMCSamplerHIP(const SMCSamplerHIP &) = delete;
    SMCSamplerHIP &operator=(const SMCSamplerHIP &) = delete;

    void resample_threshold(double t) { threshold_ = t; }
    double resample_threshold() const { return threshold_; }
    void resample(ResampleScheme s) { scheme_ = s; }

    /// all T steps, asynchronous on the current stream
    void iterate()
    {
        create();
        hip_check(mckl_hip_smc_run(s_, current_stream()));
    }

    double log_likelihood()
    {
        double ll = 0;
        hip_check(mckl_hip_smc_result(s_, &ll, nullptr, nullptr));
        return ll;
    }

    std::vector<double> ess_history()
    {
        std::vector<double> e(y_.size());
        hip_check(mckl_hip_smc_result(s_, nullptr, e.data(), nullptr));
        return e;
    }

    /// selection estimator of x^1..x^k (SMCEstimator at step 0); set before
    /// iterate()
    void estimate_moments(int k)
    {
        create();
        hip_check(mckl_hip_smc_estimate_moments(s_, k));
        k_ = k;
    }

    /// T x k estimates, row-major (one row per iteration)
    std::vector<double> estimates()
    {
        std::vector<double> e(y_.size() * static_cast<std::size_t>(k_));
        hip_check(mckl_hip_smc_estimates(s_, e.data()));
        return e;
    }

  private:
    int model_;
    std::array<double, 3> theta_;
    std::vector<double> y_;
    std::size_t N_;
    std::uint64_t seed_;
    double threshold_ = 0.5;
    int scheme_ = MCKL_HIP_SYSTEMATIC;
    int k_ = 0;
    mckl_hip_smc *s_ = nullptr;

    void create()
    {
        if (!s_)
            hip_check(mckl_hip_smc_create(model_, theta_.data(), y_.size(),
                y_.data(), N_, seed_, threshold_, scheme_, &s_));
    }
};

/// PMCMCMutation<Param, T> (algorithm/pmcmc.hpp:71-158) for the built-in
/// models, param = theta.  The particle filter is a one-filter batch (one
/// persistent launch per run) whose log-likelihood is the reference's
/// log_nc; update objects and the acceptance draw (log of
/// U01CanonicalDistribution<double>, the reference's default
/// U01Distribution) use rng().  A proposal whose prior is -inf is rejected
/// without a run.  Filter run r uses seed + 2^32 (r + 1).
class PMCMCMutationHIP
{
  public:
    using param_type = std::array<double, 3>;
    using eval_type = std::function<double(Philox4x32 &, param_type &)>;
    using prior_type = std::function<double(const param_type &)>;

    PMCMCMutationHIP(std::size_t N, std::size_t M, prior_type prior,
        int model, const std::vector<double> &y, std::uint32_t seed = 1,
        double threshold = 0.5, int scheme = MCKL_HIP_STRATIFIED)
        : prior_(std::move(prior)), rng_(seed), seed_(seed)
    {
        if (y.size() < M)
            throw RuntimeAssert("PMCMCMutationHIP: fewer than M observations");
        const std::uint32_t n = static_cast<std::uint32_t>(N);
        const double init[3] = {0.5, 1.0, 1.0};
        const std::uint64_t s = seed;
        hip_check(mckl_hip_smc_batch_create(model, 1, &n, init, M, y.data(),
            &s, threshold, scheme, &pf_));
    }
    ~PMCMCMutationHIP() { mckl_hip_smc_batch_destroy(pf_); }
    PMCMCMutationHIP(const PMCMCMutationHIP &) = delete;
    PMCMCMutationHIP &operator=(const PMCMCMutationHIP &) = delete;

    void prior(prior_type p) { prior_ = std::move(p); }
    std::size_t update(eval_type e)
    {
        eval_.push_back(std::move(e));
        return eval_.size() - 1;
    }
    void reset() { eval_.clear(); }
    Philox4x32 &rng() { return rng_; }
    double log_nc() const { return log_nc_; }

    std::size_t operator()(std::size_t iter, param_type &param)
    {
        if (iter == 0) {
            log_nc_ = loglik(param);
            return 0;
        }
        const double lnc = log_nc_;
        double prob = -prior_(param) - lnc;
        param_type p(param);
        for (auto &e : eval_)
            prob += e(rng_, p);
        prob += prior_(p);
        double lnew = -HUGE_VAL;
        if (prob > -HUGE_VAL) {
            lnew = loglik(p);
            prob += lnew;
        } else {
            ++runs_;
        }
        double *d = nullptr;
        double u = 0;
        if (hipMalloc(reinterpret_cast<void **>(&d), sizeof(double)) !=
            hipSuccess)
            throw RuntimeAssert("mckl_hip: hipMalloc failed");
        int rc = mckl_hip_rand_u01_f64(&rng_.state(), MCKL_HIP_U01_CANONICAL,
            1, d, current_stream());
        if (rc == MCKL_HIP_OK &&
            hipMemcpy(&u, d, sizeof(u), hipMemcpyDeviceToHost) != hipSuccess)
            rc = MCKL_HIP_EHIP;
        hipFree(d);
        hip_check(rc);
        if (std::log(u) < prob) {
            param = p;
            log_nc_ = lnew;
            return 1;
        }
        return 0;
    }

  private:
    prior_type prior_;
    std::vector<eval_type> eval_;
    Philox4x32 rng_;
    std::uint64_t seed_;
    std::uint64_t runs_ = 0;
    double log_nc_ = 0;
    mckl_hip_smc_batch *pf_ = nullptr;

    double loglik(const param_type &p)
    {
        const std::uint64_t s = seed_ + ((runs_ + 1) << 32);
        ++runs_;
        hip_check(mckl_hip_smc_batch_set_params(pf_, p.data(), &s));
        hip_check(mckl_hip_smc_batch_run(pf_, 1, current_stream()));
        double ll = 0;
        hip_check(mckl_hip_smc_batch_result(pf_, &ll));
        return ll;
    }
};

/// SMCEstimator::estimate's weighted sums (algorithm/smc.hpp:89-118) on
/// device arrays: d_r[j] = sum_i d_w[i] u(i, j)
inline void estimate(std::size_t n, std::size_t dim, MatrixLayout layout,
    const double *d_u, const double *d_w, double *d_r)
{
    hip_check(mckl_hip_estimate_f64(
        n, dim, static_cast<int>(layout), d_u, d_w, d_r, current_stream()));
}

inline void estimate(std::size_t n, std::size_t dim, MatrixLayout layout,
    const float *d_u, const double *d_w, double *d_r)
{
    hip_check(mckl_hip_estimate_f32(
        n, dim, static_cast<int>(layout), d_u, d_w, d_r, current_stream()));
}

} // namespace mckl

#endif
