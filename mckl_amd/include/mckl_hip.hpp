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

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <istream>
#include <limits>
#include <memory>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
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
        (void)hipFree(d);
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

/// Batched SMC for the built-in models (no reference counterpart): nfilters
/// independent filters of uneven sizes in one persistent launch.  theta: 3
/// values per filter, y: T per filter, one seed per filter; filter f gives
/// the log-likelihood SMCSamplerHIP gives with the same inputs, bit for bit.
class SMCBatchHIP
{
  public:
    SMCBatchHIP(int model, const std::vector<std::uint32_t> &N,
        const std::vector<double> &theta, std::size_t T,
        const std::vector<double> &y, const std::vector<std::uint64_t> &seed,
        double threshold = 0.5, int scheme = MCKL_HIP_STRATIFIED)
        : n_(N.size())
    {
        if (theta.size() != 3 * n_ || y.size() != T * n_ ||
            seed.size() != n_)
            throw RuntimeAssert("SMCBatchHIP: inconsistent sizes");
        hip_check(mckl_hip_smc_batch_create(model, n_, N.data(),
            theta.data(), T, y.data(), seed.data(), threshold, scheme, &b_));
    }
    ~SMCBatchHIP()
    {
        if (b_)
            mckl_hip_smc_batch_destroy(b_);
    }
    SMCBatchHIP(const SMCBatchHIP &) = delete;
    SMCBatchHIP &operator=(const SMCBatchHIP &) = delete;

    std::size_t size() const { return n_; }
    /// launch configuration (changes no result bit): workgroup width,
    /// occupancy cap, and with split_tiles > 0 a concurrent launch of
    /// small_block-thread workgroups for the filters of fewer tiles
    void set_tune(int block, int waves_per_eu = 0, int split_tiles = 0,
        int small_block = 256)
    {
        mckl_hip_tune t;
        t.block = block;
        t.waves_per_eu = waves_per_eu;
        t.split_tiles = split_tiles;
        t.small_block = small_block;
        hip_check(mckl_hip_smc_batch_set_tune(b_, &t));
    }
    /// the library's autotuner (mckl_hip_smc_batch_autotune); cache_path
    /// may be empty (no cache)
    mckl_hip_tune autotune(const std::string &cache_path = std::string(),
        double *ms = nullptr)
    {
        mckl_hip_tune t;
        hip_check(mckl_hip_smc_batch_autotune(b_,
            cache_path.empty() ? nullptr : cache_path.c_str(), &t, ms,
            current_stream()));
        return t;
    }
    /// new parameters and seeds for every filter (sizes, observations,
    /// threshold and scheme stay)
    void set_params(const std::vector<double> &theta,
        const std::vector<std::uint64_t> &seed)
    {
        if (theta.size() != 3 * n_ || seed.size() != n_)
            throw RuntimeAssert("SMCBatchHIP: inconsistent sizes");
        hip_check(mckl_hip_smc_batch_set_params(b_, theta.data(), seed.data()));
    }
    /// all filters, asynchronous on the current stream
    void run(bool use_graph = true)
    {
        hip_check(mckl_hip_smc_batch_run(b_, use_graph ? 1 : 0,
            current_stream()));
    }
    std::vector<double> log_likelihood()
    {
        std::vector<double> ll(n_);
        hip_check(mckl_hip_smc_batch_result(b_, ll.data()));
        return ll;
    }

  private:
    std::size_t n_;
    mckl_hip_smc_batch *b_ = nullptr;
};

/// Gradient of the SMC log-likelihood with respect to the built-in models'
/// theta by a hand-written adjoint (no reference counterpart): nbatch
/// points, N particles each, ancestors and decisions of the forward pass
/// held fixed.  y: T values shared by all points, or T per point.
/// replay(theta) (with tape) evaluates the differentiated function.
class SMCGradientHIP
{
  public:
    SMCGradientHIP(int model, std::size_t nbatch, std::size_t N,
        std::size_t T, const std::vector<double> &y, bool y_per_point,
        const std::vector<std::uint64_t> &seed, double threshold = 0.5,
        int scheme = MCKL_HIP_SYSTEMATIC, std::size_t K = 50,
        bool tape = false)
        : n_(nbatch)
    {
        if (y.size() != (y_per_point ? T * nbatch : T) ||
            seed.size() != nbatch)
            throw RuntimeAssert("SMCGradientHIP: inconsistent sizes");
        hip_check(mckl_hip_smc_grad_create(model, nbatch, N, T, y.data(),
            y_per_point ? 1 : 0, seed.data(), threshold, scheme, K,
            tape ? 1 : 0, &g_));
    }
    ~SMCGradientHIP()
    {
        if (g_)
            mckl_hip_smc_grad_destroy(g_);
    }
    SMCGradientHIP(const SMCGradientHIP &) = delete;
    SMCGradientHIP &operator=(const SMCGradientHIP &) = delete;

    /// log-likelihoods (nbatch) and gradients (3 per point); synchronises
    void run(const std::vector<double> &theta, std::vector<double> &loglik,
        std::vector<double> &grad)
    {
        if (theta.size() != 3 * n_)
            throw RuntimeAssert("SMCGradientHIP: theta holds 3 per point");
        loglik.resize(n_);
        grad.resize(3 * n_);
        hip_check(mckl_hip_smc_grad_run(g_, theta.data(), loglik.data(),
            grad.data(), current_stream()));
    }
    /// the differentiated function at theta (needs tape = true)
    std::vector<double> replay(const std::vector<double> &theta)
    {
        if (theta.size() != 3 * n_)
            throw RuntimeAssert("SMCGradientHIP: theta holds 3 per point");
        std::vector<double> v(n_);
        hip_check(mckl_hip_smc_grad_replay(g_, theta.data(), v.data(),
            current_stream()));
        return v;
    }

  private:
    std::size_t n_;
    mckl_hip_smc_grad *g_ = nullptr;
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
        (void)hipFree(d);
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

/// Device array owned by the host objects below (hipMalloc / hipFree)
template <typename T>
class DeviceArray
{
  public:
    DeviceArray() = default;
    explicit DeviceArray(std::size_t n) { resize(n); }
    ~DeviceArray() { release(); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : p_(o.p_), n_(o.n_)
    {
        o.p_ = nullptr;
        o.n_ = 0;
    }
    DeviceArray &operator=(DeviceArray &&o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    void resize(std::size_t n)
    {
        if (n == n_)
            return;
        release();
        void *q = nullptr;
        if (n && hipMalloc(&q, n * sizeof(T)) != hipSuccess)
            throw RuntimeAssert("mckl_hip: device allocation failed");
        p_ = static_cast<T *>(q);
        n_ = n;
    }
    T *data() const { return p_; }
    std::size_t size() const { return n_; }

  private:
    T *p_ = nullptr;
    std::size_t n_ = 0;

    void release()
    {
        if (p_)
            (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
};

inline void select_rows(std::size_t n, std::size_t dim,
    const std::uint32_t *d_idx, const float *src, float *dst)
{
    hip_check(mckl_hip_select_f32(n, dim, d_idx, src, dst, current_stream()));
}

inline void select_rows(std::size_t n, std::size_t dim,
    const std::uint32_t *d_idx, const double *src, double *dst)
{
    hip_check(mckl_hip_select_f64(n, dim, d_idx, src, dst, current_stream()));
}

inline void select_cols(std::size_t n, std::size_t N, std::size_t dim,
    const std::uint32_t *d_idx, const float *src, float *dst)
{
    hip_check(mckl_hip_select_cols_f32(
        n, N, dim, d_idx, src, dst, current_stream()));
}

inline void select_cols(std::size_t n, std::size_t N, std::size_t dim,
    const std::uint32_t *d_idx, const double *src, double *dst)
{
    hip_check(mckl_hip_select_cols_f64(
        n, N, dim, d_idx, src, dst, current_stream()));
}

/// Particle<T> (core/particle.hpp:277-360) in device memory: the state is
/// N rows of dim values of T (float or double), row-major as
/// StateMatrix<RowMajor, Dynamic, T> or, with layout ColMajor, dim columns
/// of N values as StateMatrix<ColMajor, Dynamic, T> (col_data(j) is column
/// j); weight() is the device Weight; rng() is the engine ResampleEval
/// draws from.
template <typename T, typename RNG = Philox4x32>
class Particle
{
    static_assert(std::is_same<T, float>::value ||
            std::is_same<T, double>::value,
        "Particle: the device state type is float or double");

  public:
    using size_type = std::size_t;
    using state_type = T;
    using rng_type = RNG;

    explicit Particle(std::size_t N, std::size_t dim = 1, RNG rng = RNG(),
        MatrixLayout layout = RowMajor)
        : dim_(dim), layout_(layout), rng_(rng)
    {
        if (N == 0 || dim == 0)
            throw RuntimeAssert("Particle: N and dim must be positive");
        alloc(N);
    }
    Particle(const Particle &) = delete;
    Particle &operator=(const Particle &) = delete;

    size_type size() const { return n_; }
    std::size_t dim() const { return dim_; }
    MatrixLayout layout() const { return layout_; }
    /// device pointer to the N x dim state in layout()
    T *state() { return x_.data(); }
    const T *state() const { return x_.data(); }
    /// device pointer to column j (ColMajor layout only)
    T *col_data(std::size_t j)
    {
        if (layout_ != ColMajor || j >= dim_)
            throw RuntimeAssert(
                "Particle::col_data: ColMajor layout and j < dim() required");
        return x_.data() + j * n_;
    }
    Weight &weight() { return weight_; }
    const Weight &weight() const { return weight_; }
    RNG &rng() { return rng_; }

    /// Particle::select (core/particle.hpp:324-331): row j of the new state
    /// is row d_idx[j] of the old one (j < n; d_idx a device array of row
    /// numbers below size()), then equal weights of size n.  With validate
    /// (the default) the row numbers are read back and checked first, and a
    /// number not below size() throws RuntimeAssert before the gather;
    /// ResampleEval, whose indices are parents by construction, skips it.
    void select(size_type n, const std::uint32_t *d_idx, bool validate = true)
    {
        if (validate && n > 0) {
            std::vector<std::uint32_t> h(n);
            const auto st = static_cast<hipStream_t>(current_stream());
            if (hipMemcpyAsync(h.data(), d_idx, n * sizeof(std::uint32_t),
                    hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                throw RuntimeAssert(
                    "Particle::select: reading the row numbers failed");
            for (std::uint32_t v : h)
                if (v >= n_)
                    throw RuntimeAssert(
                        "Particle::select: row numbers must lie below size()");
        }
        if (n == n_) {
            gather(n, d_idx, spare_.data());
            std::swap(x_, spare_);
            weight_.set_equal();
            return;
        }
        DeviceArray<T> nx(n * dim_);
        gather(n, d_idx, nx.data());
        x_ = std::move(nx);
        spare_.resize(n * dim_);
        w_.resize(n);
        n_ = n;
        weight_ = Weight(n, w_.data(), stats_.data());
    }

  private:
    std::size_t n_ = 0;
    std::size_t dim_;
    MatrixLayout layout_;
    RNG rng_;
    DeviceArray<T> x_, spare_;

    void gather(size_type n, const std::uint32_t *d_idx, T *dst)
    {
        if (layout_ == ColMajor)
            select_cols(n, n_, dim_, d_idx, x_.data(), dst);
        else
            select_rows(n, dim_, d_idx, x_.data(), dst);
    }
    DeviceArray<double> w_, stats_;
    Weight weight_{0, nullptr, nullptr};

    void alloc(std::size_t N)
    {
        x_.resize(N * dim_);
        spare_.resize(N * dim_);
        w_.resize(N);
        stats_.resize(2);
        n_ = N;
        if (hipMemset(x_.data(), 0, N * dim_ * sizeof(T)) != hipSuccess)
            throw RuntimeAssert("mckl_hip: clearing the state failed");
        weight_ = Weight(N, w_.data(), stats_.data());
    }
};

/// ResampleEval<T> (algorithm/resample.hpp:470-511) of a built-in scheme:
/// replication numbers from Particle::rng() and the weights, the stable
/// rep -> index transform, Particle::select
template <typename T, typename RNG = Philox4x32>
class ResampleEval
{
  public:
    explicit ResampleEval(ResampleScheme scheme)
        : scheme_(scheme), buf_(std::make_shared<Buffers>())
    {
    }

    void operator()(std::size_t, Particle<T, RNG> &particle) const
    {
        const std::size_t N = particle.size();
        buf_->rep.resize(N);
        buf_->idx.resize(N);
        hip_check(mckl_hip_resample(scheme_, N, N, &particle.rng().state(),
            particle.weight().data(), buf_->rep.data(), buf_->idx.data(),
            current_stream()));
        particle.select(N, buf_->idx.data(), false);
    }

  private:
    struct Buffers {
        DeviceArray<std::uint32_t> rep, idx;
    };
    int scheme_;
    std::shared_ptr<Buffers> buf_;
};

/// SMCEstimator<T, double> (algorithm/smc.hpp:48-136).  eval(iter, dim,
/// particle, d_u) fills the device array d_u: N x dim values (RowMajor),
/// dim x N (ColMajor), or the dim values of the estimate itself
/// (record_only).  Otherwise estimate() stores r_j = sum_i W_i u(i, j),
/// each sum exactly rounded (the reference calls an external BLAS dgemv).
/// Estimates are kept on the host, one row per call.
template <typename T, typename RNG = Philox4x32>
class SMCEstimator
{
  public:
    using eval_type = std::function<void(
        std::size_t, std::size_t, Particle<T, RNG> &, double *)>;

    SMCEstimator(std::size_t dim, eval_type eval,
        MatrixLayout layout = MatrixLayout::RowMajor, bool record_only = false)
        : dim_(dim), eval_(std::move(eval)), layout_(layout),
          record_only_(record_only), buf_(std::make_shared<Buffers>())
    {
        if (dim == 0)
            throw RuntimeAssert("SMCEstimator: dim must be positive");
    }

    std::size_t dim() const { return dim_; }
    std::size_t num_iter() const { return rows_.size() / dim_; }
    bool record_only() const { return record_only_; }
    void clear() { rows_.clear(); }
    void reserve(std::size_t n) { rows_.reserve(rows_.size() + n * dim_); }

    /// SMCEstimator::estimate(eval, layout, record_only) (smc.hpp:76-84)
    void estimate(eval_type eval,
        MatrixLayout layout = MatrixLayout::RowMajor, bool record_only = false)
    {
        eval_ = std::move(eval);
        layout_ = layout;
        record_only_ = record_only;
    }

    /// SMCEstimator::estimate(iter, particle) (smc.hpp:86-118); reads the
    /// result to the host (synchronises the stream)
    void estimate(std::size_t iter, Particle<T, RNG> &particle)
    {
        if (!eval_)
            throw RuntimeAssert(
                "**Estimator::eval** used with an invalid evaluation object");
        const std::size_t n = particle.size();
        buf_->r.resize(dim_);
        if (record_only_) {
            eval_(iter, dim_, particle, buf_->r.data());
        } else {
            buf_->u.resize(n * dim_);
            eval_(iter, dim_, particle, buf_->u.data());
            mckl::estimate(n, dim_, layout_, buf_->u.data(),
                particle.weight().data(), buf_->r.data());
        }
        const std::size_t k = rows_.size();
        rows_.resize(k + dim_);
        if (hipStreamSynchronize(static_cast<hipStream_t>(current_stream())) !=
                hipSuccess ||
            hipMemcpy(rows_.data() + k, buf_->r.data(), dim_ * sizeof(double),
                hipMemcpyDeviceToHost) != hipSuccess)
            throw RuntimeAssert("mckl_hip: reading an estimate failed");
    }

    /// row i (iteration i) of the estimates
    const double *row_data(std::size_t i) const
    {
        return rows_.data() + i * dim_;
    }

    /// variable j over the iterations (EstimateMatrix::variable)
    template <typename OutputIter>
    OutputIter read_estimate(std::size_t j, OutputIter first) const
    {
        for (std::size_t i = 0; i != num_iter(); ++i)
            *first++ = rows_[i * dim_ + j];
        return first;
    }

    /// Estimator::average (core/estimator.hpp:62-93): rows cut, cut + thin,
    /// ... summed and, as the reference does, divided by the number of
    /// stored iterations; row cut itself when fewer than thin rows follow
    template <typename OutputIter>
    OutputIter average(
        OutputIter first, std::size_t cut = 0, std::size_t thin = 1) const
    {
        const std::size_t n = num_iter();
        if (cut >= n)
            return first;
        thin = thin < 1 ? 1 : thin;
        if (n - cut < thin)
            return std::copy(row_data(cut), row_data(cut) + dim_, first);
        std::vector<double> sum(dim_, 0.0);
        for (; cut < n; cut += thin)
            for (std::size_t j = 0; j != dim_; ++j)
                sum[j] += rows_[cut * dim_ + j];
        for (auto &v : sum)
            v *= 1.0 / static_cast<double>(n);
        return std::copy(sum.begin(), sum.end(), first);
    }

  private:
    struct Buffers {
        DeviceArray<double> u, r;
    };
    std::size_t dim_;
    eval_type eval_;
    MatrixLayout layout_;
    bool record_only_;
    std::shared_ptr<Buffers> buf_;
    std::vector<double> rows_;
};

/// SMCSampler<T> (algorithm/smc.hpp:151-421) on a device Particle<T>.
/// Evaluation objects eval(iter, particle) enqueue their own device work on
/// current_stream().  iterate(n) runs do_iterate (smc.hpp:389-406) n times:
/// selection evals, selection estimators, the size and ESS recorded (the
/// ESS read to the host), resample evals when ESS < N * threshold,
/// resample estimators, mutation evals, mutation estimators.
template <typename T, typename RNG = Philox4x32>
class SMCSampler
{
  public:
    using state_type = T;
    using size_type = std::size_t;
    using eval_type = std::function<void(std::size_t, Particle<T, RNG> &)>;
    using estimator_type = SMCEstimator<T, RNG>;

    explicit SMCSampler(std::size_t N, std::size_t dim = 1, RNG rng = RNG(),
        MatrixLayout layout = RowMajor)
        : particle_(N, dim, rng, layout)
    {
    }

    size_type size() const { return particle_.size(); }
    std::size_t num_iter() const { return size_history_.size(); }
    void reserve(std::size_t n)
    {
        for (auto &slot : est_)
            for (auto &e : slot)
                e.reserve(n);
        size_history_.reserve(size_history_.size() + n);
        ess_history_.reserve(ess_history_.size() + n);
    }

    /// remove every evaluation object and estimator, clear the history
    void reset()
    {
        for (auto &slot : eval_)
            slot.clear();
        for (auto &slot : est_)
            slot.clear();
        clear();
    }

    /// clear the history (smc.hpp:205-212)
    void clear()
    {
        for (auto &slot : est_)
            for (auto &e : slot)
                e.clear();
        iter_ = 0;
        size_history_.clear();
        ess_history_.clear();
    }

    double resample_threshold() const { return threshold_; }
    void resample_threshold(double t) { threshold_ = t; }
    static double resample_threshold_never()
    {
        return -std::numeric_limits<double>::infinity();
    }
    static double resample_threshold_always()
    {
        return std::numeric_limits<double>::infinity();
    }

    std::size_t selection(eval_type e) { return add(0, std::move(e)); }
    std::size_t resample(eval_type e) { return add(1, std::move(e)); }
    /// the built-in ResampleEval of a scheme (smc.hpp:246-270)
    std::size_t resample(ResampleScheme scheme)
    {
        return add(1, ResampleEval<T, RNG>(scheme));
    }
    std::size_t mutation(eval_type e) { return add(2, std::move(e)); }
    eval_type &selection(std::size_t k) { return eval_[0].at(k); }
    eval_type &resample(std::size_t k) { return eval_[1].at(k); }
    eval_type &mutation(std::size_t k) { return eval_[2].at(k); }

    std::size_t selection_estimator(estimator_type e)
    {
        return add_est(0, std::move(e));
    }
    std::size_t resample_estimator(estimator_type e)
    {
        return add_est(1, std::move(e));
    }
    std::size_t mutation_estimator(estimator_type e)
    {
        return add_est(2, std::move(e));
    }
    estimator_type &selection_estimator(std::size_t k)
    {
        return est_[0].at(k);
    }
    estimator_type &resample_estimator(std::size_t k)
    {
        return est_[1].at(k);
    }
    estimator_type &mutation_estimator(std::size_t k)
    {
        return est_[2].at(k);
    }

    void iterate(std::size_t n = 1)
    {
        if (n > 1)
            reserve(n);
        for (std::size_t i = 0; i != n; ++i)
            do_iterate();
    }

    Particle<T, RNG> &particle() { return particle_; }

    template <typename OutputIter>
    OutputIter read_size_history(OutputIter first) const
    {
        return std::copy(size_history_.begin(), size_history_.end(), first);
    }

    template <typename OutputIter>
    OutputIter read_ess_history(OutputIter first) const
    {
        return std::copy(ess_history_.begin(), ess_history_.end(), first);
    }

    /// Sampler::summary (core/sampler.hpp:80-118): every estimator's
    /// estimates side by side, num_iter() rows, in the given layout
    std::vector<double> summary(
        MatrixLayout layout = MatrixLayout::RowMajor) const
    {
        std::size_t ncol = 0;
        for (auto &slot : est_)
            for (auto &e : slot)
                ncol += e.dim();
        const std::size_t nrow = num_iter();
        std::vector<double> m(nrow * ncol);
        std::size_t c = 0;
        for (auto &slot : est_) {
            for (auto &e : slot) {
                for (std::size_t i = 0; i != nrow && i != e.num_iter(); ++i)
                    for (std::size_t j = 0; j != e.dim(); ++j) {
                        const double v = e.row_data(i)[j];
                        if (layout == MatrixLayout::RowMajor)
                            m[i * ncol + c + j] = v;
                        else
                            m[(c + j) * nrow + i] = v;
                    }
                c += e.dim();
            }
        }
        return m;
    }

  private:
    Particle<T, RNG> particle_;
    std::vector<eval_type> eval_[3];
    std::vector<estimator_type> est_[3];
    std::size_t iter_ = 0;
    double threshold_ = resample_threshold_never();
    std::vector<size_type> size_history_;
    std::vector<double> ess_history_;

    std::size_t add(int slot, eval_type e)
    {
        if (!e)
            throw RuntimeAssert("SMCSampler: invalid evaluation object");
        eval_[slot].push_back(std::move(e));
        return eval_[slot].size() - 1;
    }

    std::size_t add_est(int slot, estimator_type e)
    {
        est_[slot].push_back(std::move(e));
        return est_[slot].size() - 1;
    }

    void do_iterate()
    {
        for (auto &e : eval_[0])
            e(iter_, particle_);
        for (auto &e : est_[0])
            e.estimate(iter_, particle_);
        size_history_.push_back(size());
        ess_history_.push_back(particle_.weight().ess());
        if (ess_history_.back() <
            static_cast<double>(size()) * threshold_) {
            for (auto &e : eval_[1])
                e(iter_, particle_);
        }
        for (auto &e : est_[1])
            e.estimate(iter_, particle_);
        for (auto &e : eval_[2])
            e(iter_, particle_);
        for (auto &e : est_[2])
            e.estimate(iter_, particle_);
        ++iter_;
    }
};

/// PMCMCStateMatrix's parameter and log normalising constant
/// (algorithm/pmcmc.hpp:46-66), kept beside the device particles: the
/// filter's evaluation objects read param() and call add_log_nc()
template <typename Param>
class PMCMCState
{
  public:
    using param_type = Param;
    double log_nc() const { return log_nc_; }
    void add_log_nc(double nc) { log_nc_ += nc; }
    const param_type &param() const { return param_; }
    void reset(const param_type &p, double nc = 0)
    {
        param_ = p;
        log_nc_ = nc;
    }

  private:
    param_type param_{};
    double log_nc_ = 0;
};

/// PMCMCMutation<Param, T> (algorithm/pmcmc.hpp:71-158) over SMCSampler<T>
/// with user evaluation objects, step for step: iteration 0 runs the filter
/// at the starting parameter; afterwards the proposal's filter always runs
/// (the prior is added, not tested), the acceptance draw is
/// log(U01Canonical<double>) from the particle's engine, and a rejected
/// proposal restores the parameter and log normalising constant but keeps
/// the particles.  state() is the PMCMCState the evaluation objects use.
template <typename Param, typename T, typename RNG = Philox4x32>
class PMCMCMutation
{
  public:
    using param_type = Param;
    using eval_type = std::function<double(RNG &, param_type &)>;
    using prior_type = std::function<double(const param_type &)>;
    using pf_type = SMCSampler<T, RNG>;

    PMCMCMutation(std::size_t N, std::size_t M, prior_type prior,
        std::size_t dim = 1, RNG rng = RNG())
        : M_(M), prior_(std::move(prior)), pf_(N, dim, rng), u_(1)
    {
    }

    pf_type &pf() { return pf_; }
    PMCMCState<Param> &state() { return state_; }
    void prior(prior_type p) { prior_ = std::move(p); }
    std::size_t update(eval_type e)
    {
        eval_.push_back(std::move(e));
        return eval_.size() - 1;
    }
    void reset()
    {
        eval_.clear();
        pf_.reset();
    }

    std::size_t operator()(std::size_t iter, param_type &param)
    {
        if (iter == 0) {
            run(param);
            return 0;
        }
        const double lnc = state_.log_nc();
        double prob = -prior_(param) - lnc;
        param_type p(param);
        for (auto &e : eval_)
            prob += e(pf_.particle().rng(), p);
        prob += prior_(p);
        run(p);
        prob += state_.log_nc();
        hip_check(mckl_hip_rand_u01_f64(&pf_.particle().rng().state(),
            MCKL_HIP_U01_CANONICAL, 1, u_.data(), current_stream()));
        double u = 0;
        if (hipStreamSynchronize(static_cast<hipStream_t>(current_stream())) !=
                hipSuccess ||
            hipMemcpy(&u, u_.data(), sizeof(u), hipMemcpyDeviceToHost) !=
                hipSuccess)
            throw RuntimeAssert("mckl_hip: reading the acceptance draw failed");
        if (std::log(u) < prob) {
            param = std::move(p);
            return 1;
        }
        state_.reset(param, lnc);
        return 0;
    }

  private:
    std::size_t M_;
    prior_type prior_;
    std::vector<eval_type> eval_;
    pf_type pf_;
    PMCMCState<Param> state_;
    DeviceArray<double> u_;

    void run(const param_type &p)
    {
        pf_.clear();
        state_.reset(p, 0);
        pf_.iterate(M_);
    }
};

} // namespace mckl

#endif
