// ref_smc.cpp -- the reference's SMCSampler<T> driven on a table-noise model.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
// the reference headers where they lie, with the flags of ref_driver.cpp.
// A library of its own because SMCSampler::iterate instantiates
// SMCEstimator::estimate, whose dgemv is an external BLAS routine
// (internal/cblas.hpp:108-118) that this image does not have: the symbol
// stays unresolved, the library is loaded with lazy binding, and no
// estimator is ever attached, so dgemv is never called.

#include <mckl/algorithm/smc.hpp>
#include <mckl/core/state_matrix.hpp>

#include <algorithm>
#include <cstdint>
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

} // extern "C"
