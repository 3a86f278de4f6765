// ref_smc.cpp -- the reference's SMCSampler<T> driven on a table-noise model.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
// the reference headers where they lie, with the flags of ref_driver.cpp.
// A library of its own because SMCSampler::iterate instantiates
// SMCEstimator::estimate, whose dgemv is an external BLAS routine
// (internal/cblas.hpp:108-118) that this image does not have: the symbol
// stays unresolved, the library is loaded with lazy binding, and no
// estimator is ever attached, so dgemv is never called.

#include <mckl/algorithm/pmcmc.hpp>
#include <mckl/algorithm/smc.hpp>
#include <mckl/core/state_matrix.hpp>
#include <mckl/random/normal_distribution.hpp>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <type_traits>
#include <vector>

extern "C" {

// SMCSampler<StateMatrix<RowMajor, double, 1>> (algorithm/smc.hpp:151-421)
// on a linear-Gaussian model whose noise comes from a table, so that any
// implementation can replay it.  Selection eval at iteration t: x_i = z[i]
// (t = 0) or a x_i + sx z[t N + i] (t > 0), then Weight::add_log of
// v_i = -0.5 d_i d_i, d_i = (y_t - x_i) / sy.  Mutation eval: x_i = mut x_i
// (mut = 1: no mutation eval).  resample(scheme), resample_threshold(thr),
// Particle::rng() keyed by {key[0], key[1]}.  Outputs: the ESS history
// ess[T], the size history size[T] and the final state x[N].
int ref_smc_lg(std::size_t N, std::size_t T, int scheme, double threshold,
    const std::uint64_t *key, double a, double sx, double sy, double mut,
    const double *y, const double *z, double *ess, std::uint64_t *size,
    double *x)
{
    using State = mckl::StateMatrix<mckl::RowMajor, double, 1>;
    static_assert(std::is_same<mckl::Particle<State>::rng_type,
                      mckl::Philox4x32>::value,
        "the reference build here uses Philox4x32 as Particle::rng_type");
    mckl::SMCSampler<State> s(N);
    mckl::Philox4x32::key_type k;
    k[0] = static_cast<std::uint32_t>(key[0]);
    k[1] = static_cast<std::uint32_t>(key[1]);
    s.particle().rng().seed(k);
    std::vector<double> v(N);
    s.selection([&](std::size_t t, mckl::Particle<State> &p) {
        double *xs = p.state().data();
        for (std::size_t i = 0; i != N; ++i) {
            xs[i] = t == 0 ? z[i] : a * xs[i] + sx * z[t * N + i];
            const double d = (y[t] - xs[i]) / sy;
            v[i] = -0.5 * (d * d);
        }
        p.weight().add_log(v.data());
    });
    s.resample(static_cast<mckl::ResampleScheme>(scheme));
    if (mut != 1.0) {
        s.mutation([&](std::size_t, mckl::Particle<State> &p) {
            double *xs = p.state().data();
            for (std::size_t i = 0; i != N; ++i)
                xs[i] = mut * xs[i];
        });
    }
    s.resample_threshold(threshold);
    s.iterate(T);
    s.read_ess_history(ess);
    std::vector<std::size_t> sz(T);
    s.read_size_history(sz.data());
    std::copy(sz.begin(), sz.end(), size);
    std::copy_n(s.particle().state().data(), N, x);
    return 0;
}

// PMCMCMutation<std::array<double, 3>, PMCMCStateMatrix<...>>
// (algorithm/pmcmc.hpp:71-158) on the table-noise model above, theta =
// (a, sx, sy) taken from state().param().  The selection eval also adds
// log(sum_i W_i exp(v_i)) (W before the increment, summed in order, std::exp
// and std::log) to state().add_log_nc.  Update eval: a Gaussian random walk,
// p_k += scale_k z_k with z from rand(rng, NormalDistribution<double>(0, 1),
// 3, z) on the mutation's rng (Particle::rng()).  Prior: 0 inside
// |a| < 1, sx > 0, sy > 0, -inf outside.  resample(scheme) with
// resample_threshold(threshold).  Outputs per iteration: accepted flag, the
// parameter after the step and the state's log_nc.
int ref_pmcmc_lg(std::size_t N, std::size_t T, std::size_t iters, int scheme,
    double threshold, const std::uint64_t *key, const double *scale,
    const double *theta0, const double *y, const double *z, int *acc,
    double *param_out, double *lognc_out)
{
    using Param = std::array<double, 3>;
    using State = mckl::PMCMCStateMatrix<Param, mckl::RowMajor, double, 1>;
    auto prior = [](const Param &p) {
        return std::fabs(p[0]) < 1 && p[1] > 0 && p[2] > 0 ?
            0.0 :
            -std::numeric_limits<double>::infinity();
    };
    mckl::PMCMCMutation<Param, State> mut(N, T, prior);
    mckl::Philox4x32::key_type k;
    k[0] = static_cast<std::uint32_t>(key[0]);
    k[1] = static_cast<std::uint32_t>(key[1]);
    mut.pf().particle().rng().seed(k);
    std::vector<double> v(N);
    mut.pf().selection([&](std::size_t t, mckl::Particle<State> &p) {
        const Param th = p.state().param();
        double *xs = p.state().data();
        const double *W = p.weight().data();
        double s = 0;
        for (std::size_t i = 0; i != N; ++i) {
            xs[i] = t == 0 ? z[i] : th[0] * xs[i] + th[1] * z[t * N + i];
            const double d = (y[t] - xs[i]) / th[2];
            v[i] = -0.5 * (d * d);
            s += W[i] * std::exp(v[i]);
        }
        p.state().add_log_nc(std::log(s));
        p.weight().add_log(v.data());
    });
    mut.pf().resample(static_cast<mckl::ResampleScheme>(scheme));
    mut.pf().resample_threshold(threshold);
    mut.update([&](mckl::Philox4x32 &rng, Param &p) {
        mckl::NormalDistribution<double> normal(0, 1);
        double zz[3];
        mckl::rand(rng, normal, 3, zz);
        for (int j = 0; j != 3; ++j)
            p[j] += scale[j] * zz[j];
        return 0.0;
    });
    Param param{{theta0[0], theta0[1], theta0[2]}};
    for (std::size_t it = 0; it != iters; ++it) {
        acc[it] = static_cast<int>(mut(it, param));
        for (int j = 0; j != 3; ++j)
            param_out[3 * it + j] = param[j];
        lognc_out[it] = mut.pf().particle().state().log_nc();
    }
    return 0;
}

} // extern "C"
