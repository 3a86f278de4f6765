// mckl_hip_fused.hpp -- the fused device SMC samplers on a model of the
// caller's own.  The reference's sampler takes the model as an evaluation
// object whose eval_first / eval_each run on the host per step and per
// particle (smp/backend_base.hpp:113-183); here the model is a device type
// M that the sampler kernels are compiled against, so a program including
// this header must be compiled with hipcc for gfx950 (and link
// libmckl_hip.so for the error messages).
//
// The model concept (smc_tiles.hpp states it with the built-in models):
//
//   struct MyModel {                      // plain values, device pointers
//       struct Step {
//           __device__ float move(float x, float e) const;   // x_t
//           __device__ float log_weight(float x) const;      // log g(y_t|x_t)
//       };
//       __device__ float init(float e) const;                // x_0
//       __device__ Step at(uint32_t t) const;                // step t = 1..T
//   };
//
// or, for a state of D > 1 floats,
//
//   struct MyModel {
//       static constexpr int dim = D;
//       struct Step {
//           __device__ void move(const float *xp, const float *e,
//               float *xn) const;                            // D values each
//           __device__ float log_weight(const float *x) const;
//       };
//       __device__ void init(const float *e, float *x) const;
//       __device__ Step at(uint32_t t) const;
//   };
//
// e is a standard normal draw from the samplers' counter-based stream
// (particle i, step t, component d: the stream's normal i D + d; the same
// for every model), states are FP32, weights follow the bootstrap filter:
// after the move, lw_t = carry + log_weight.  The moment estimates of a
// vector state are those of its first component.
// The samplers are the library's own (plain: two launches per step captured
// into one HIP graph; batched: one persistent launch for all filters), with
// all six resampling schemes; a model that reproduces the built-in
// linear-Gaussian or stochastic-volatility arithmetic gives the built-in
// samplers' bits.
//
//   mckl::FusedSMCSampler<M>  one filter of N particles over T steps
//   mckl::FusedSMCBatch<M>    many independent filters of uneven sizes in
//                             one launch (a fixed 256-thread workgroup per
//                             filter; the autotuned form is the C ABI's, on
//                             the built-in models)
#ifndef MCKL_HIP_FUSED_HPP
#define MCKL_HIP_FUSED_HPP

#include "mckl_hip.hpp"

#include "../csrc/smc_plain.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

namespace mckl {

/// SMCSampler<T>::iterate (algorithm/smc.hpp:352-406) for a bootstrap
/// filter whose move and weight are the device model M; iterate() runs all
/// T steps on the device.
template <typename M>
class FusedSMCSampler
{
  public:
    FusedSMCSampler(const M &model, std::size_t N, std::size_t T,
        std::uint64_t seed, double resample_threshold = 0.5,
        int scheme = MCKL_HIP_SYSTEMATIC)
    {
        hip_check(mckl_hip::plain_init(
            &s_, model, N, T, seed, resample_threshold, scheme));
    }

    FusedSMCSampler(const FusedSMCSampler &) = delete;
    FusedSMCSampler &operator=(const FusedSMCSampler &) = delete;

    ~FusedSMCSampler() { mckl_hip::plain_release(&s_); }

    std::size_t size() const { return s_.A.N; }
    std::size_t iterations() const { return s_.T; }

    /// the whole T-step run, enqueued on `stream` (no host synchronisation)
    void iterate(hipStream_t stream = nullptr)
    {
        hip_check(mckl_hip::plain_run(&s_, stream));
    }

    /// log Z-hat of the last run (waits for it)
    double log_likelihood()
    {
        double ll = 0.0;
        hip_check(mckl_hip::plain_result(&s_, &ll, nullptr, nullptr));
        return ll;
    }

    /// ESS after each step's weighting (read_ess_history)
    std::vector<double> ess_history()
    {
        std::vector<double> v(s_.T);
        hip_check(mckl_hip::plain_result(&s_, nullptr, v.data(), nullptr));
        return v;
    }

    /// whether each step resampled (read_size_history's resampling flags)
    std::vector<std::int32_t> resampled_history()
    {
        std::vector<std::int32_t> v(s_.T);
        hip_check(mckl_hip::plain_result(&s_, nullptr, nullptr, v.data()));
        return v;
    }

    /// SMCEstimator of the moments x^1..x^k at every step (k = 0: none);
    /// takes effect at the next iterate()
    void estimate_moments(int k)
    {
        hip_check(mckl_hip::plain_set_moments(&s_, k));
    }

    /// T x k estimates of the last run, step-major
    std::vector<double> estimates()
    {
        std::vector<double> v(s_.T * (std::size_t)s_.kmom);
        hip_check(mckl_hip::plain_estimates(&s_, v.data()));
        return v;
    }

  private:
    mckl_hip::PlainSampler<M> s_;
};

/// Independent bootstrap filters with per-filter models, particle counts
/// (at most 131072 each) and RNG keys, all run in one persistent launch.
template <typename M>
class FusedSMCBatch
{
  public:
    FusedSMCBatch(const std::vector<M> &models,
        const std::vector<std::uint32_t> &N, std::size_t T,
        const std::vector<std::uint64_t> &seeds,
        double resample_threshold = 0.5, int scheme = MCKL_HIP_STRATIFIED)
        : nf_(models.size()), T_(T)
    {
        using namespace mckl_hip;
        if (nf_ == 0 || N.size() != nf_ || seeds.size() != nf_ ||
            nf_ >= (1u << 31))
            throw RuntimeAssert("FusedSMCBatch: inconsistent sizes");
        std::vector<FilterDescT<M>> fd(nf_);
        std::uint64_t total = 0;
        for (std::size_t f = 0; f < nf_; ++f) {
            hip_check(check_common(N[f], T, resample_threshold, scheme));
            if (N[f] > (std::size_t)kMaxTiles * kTile)
                throw RuntimeAssert(
                    "FusedSMCBatch: a filter holds at most 131072 particles");
            SmcArgs &A = fd[f].A;
            std::memset(&A, 0, sizeof(A));
            A.k0 = (std::uint32_t)seeds[f];
            A.k1 = (std::uint32_t)(seeds[f] >> 32);
            A.N = N[f];
            A.ntiles = (std::uint32_t)ceil_div(N[f], kTile);
            A.tauN = resample_threshold * (double)N[f];
            A.scheme = scheme;
            A.T = (std::uint32_t)T;
            fd[f].model = models[f];
            fd[f].off = total;
            total += (std::uint64_t)A.ntiles * kTile;
        }
        // largest filters first (longest-processing-time order)
        std::vector<std::uint32_t> order(nf_);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(),
            [&](std::uint32_t a, std::uint32_t b) { return N[a] > N[b]; });
        try {
            alloc(fd_, nf_);
            alloc(order_, nf_);
            alloc(queue_, 2);
            alloc(xa_, total * (std::size_t)model_dim<M>::value);
            alloc(xb_, total * (std::size_t)model_dim<M>::value);
            alloc(lw_, total);
            alloc(idx_, total);
            alloc(cnt_, total);
            alloc(E_, total);
            alloc(zpos_, total);
            if (scheme == MCKL_HIP_MULTINOMIAL || scheme == MCKL_HIP_RESIDUAL)
                alloc(u_, total);
            alloc(logz_, nf_);
            hip_check(hip_status(hipMemcpy(fd_, fd.data(),
                nf_ * sizeof(FilterDescT<M>), hipMemcpyHostToDevice),
                "FusedSMCBatch"));
            hip_check(hip_status(hipMemcpy(order_, order.data(),
                nf_ * sizeof(std::uint32_t), hipMemcpyHostToDevice),
                "FusedSMCBatch"));
            int dev = 0, ncu = 0, occ = 0;
            hip_check(hip_status(hipGetDevice(&dev), "hipGetDevice"));
            hip_check(hip_status(hipDeviceGetAttribute(&ncu,
                hipDeviceAttributeMultiprocessorCount, dev), "attribute"));
            hip_check(hip_status(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, k_batch<1, M>, kGroup, 0), "occupancy"));
            grid_ = (std::uint32_t)std::min<std::size_t>(
                (std::size_t)ncu * (std::size_t)std::max(1, occ), nf_);
            hip_check(hip_status(
                hipEventCreateWithFlags(&done_, hipEventDisableTiming),
                "FusedSMCBatch"));
        } catch (...) {
            release();
            throw;
        }
    }

    FusedSMCBatch(const FusedSMCBatch &) = delete;
    FusedSMCBatch &operator=(const FusedSMCBatch &) = delete;

    ~FusedSMCBatch() { release(); }

    /// one run of every filter, enqueued on `stream`
    void iterate(hipStream_t stream = nullptr)
    {
        using namespace mckl_hip;
        k_zero_u32<<<1, 64, 0, stream>>>(queue_, 2);
        k_batch<1, M><<<grid_, kGroup, 0, stream>>>(fd_, order_,
            (std::uint32_t)nf_, queue_, xa_, xb_, lw_, idx_, cnt_, E_, zpos_,
            u_, logz_);
        hip_check(hip_status(hipGetLastError(), "k_batch"));
        hip_check(hip_status(hipEventRecord(done_, stream), "FusedSMCBatch"));
    }

    /// log Z-hat of every filter from the last run (waits for it)
    std::vector<double> log_likelihood()
    {
        std::vector<double> v(nf_);
        hip_check(mckl_hip::hip_status(hipEventSynchronize(done_),
            "FusedSMCBatch"));
        hip_check(mckl_hip::hip_status(hipMemcpy(v.data(), logz_,
            nf_ * sizeof(double), hipMemcpyDeviceToHost), "FusedSMCBatch"));
        return v;
    }

  private:
    template <typename T>
    static void alloc(T *&p, std::size_t n)
    {
        hip_check(mckl_hip::dev_alloc(&p, n));
    }

    void release()
    {
        if (done_)
            (void)hipEventDestroy(done_);
        void *ptrs[] = {fd_, order_, queue_, xa_, xb_, lw_, idx_, cnt_, E_,
            zpos_, u_, logz_};
        for (void *p : ptrs)
            if (p)
                (void)hipFree(p);
        done_ = nullptr;
        fd_ = nullptr;
    }

    std::size_t nf_ = 0, T_ = 0;
    std::uint32_t grid_ = 0;
    mckl_hip::FilterDescT<M> *fd_ = nullptr;
    std::uint32_t *order_ = nullptr, *queue_ = nullptr;
    float *xa_ = nullptr, *xb_ = nullptr, *lw_ = nullptr;
    std::uint32_t *idx_ = nullptr, *cnt_ = nullptr, *E_ = nullptr,
                  *zpos_ = nullptr;
    double *u_ = nullptr, *logz_ = nullptr;
    hipEvent_t done_ = nullptr;
};

} // namespace mckl

#endif
