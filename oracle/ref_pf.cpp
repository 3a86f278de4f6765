// ref_pf.cpp -- the reference's own particle-filter example
// (example/algorithm/include/algorithm_pf.hpp:38-150) run with BackendSEQ.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
// the reference's headers and the example's header where they lie, with the
// pinned flags of ref_driver.cpp, and loaded with lazy binding like
// libmckl_ref_smc.so: SMCSampler::iterate instantiates the estimators'
// external BLAS dgemv, and no estimator is attached here, so it is never
// called (the example's own driver attaches three; their estimates are not
// part of these fixtures).

#include "algorithm_pf.hpp"

#include <unistd.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

extern "C" {

// Seed<Philox4x32>::instance().set(seed), then the example's sampler
// (algorithm_pf.hpp:176-190 without estimators): SMCSampler<AlgorithmPF>(N),
// selection AlgorithmPFSelection<BackendSEQ>, Stratified resampling at
// threshold 0.5, iterate(min(tmax, n)) over the observations that
// AlgorithmPF reads from algorithm_pf.data in data_dir.  Outputs: *T (the
// iterations run), ess[T] and size[T] (the sampler's histories), the final
// state (4 x N column-major: p, q, s, t), the final normalised weights [N],
// and keys[4]: the key words of the RNG set's engine 0 and of
// Particle::rng() as constructed.  Returns -1 if the data file cannot be
// read.
int ref_pf(std::size_t N, std::uint64_t seed, const char *data_dir,
    std::size_t tmax, std::size_t *T, double *ess, std::uint64_t *size,
    double *state, double *weight, std::uint64_t *keys)
{
    static_assert(std::is_same<mckl::Particle<AlgorithmPF>::rng_set_type,
                      mckl::RNGSetVector<mckl::Philox4x32>>::value,
        "the reference build here uses RNGSetVector<Philox4x32>");
    char cwd[4096];
    if (getcwd(cwd, sizeof(cwd)) == nullptr || chdir(data_dir) != 0)
        return -1;
    mckl::Seed<mckl::Philox4x32>::instance().set(seed);
    mckl::SMCSampler<AlgorithmPF> sampler(N);
    const int back = chdir(cwd);
    if (back != 0 || sampler.particle().state().n() == 0)
        return -1;
    const auto k0 = sampler.particle().rng(0).key();
    const auto kr = sampler.particle().rng().key();
    keys[0] = k0[0];
    keys[1] = k0[1];
    keys[2] = kr[0];
    keys[3] = kr[1];
    sampler.selection(AlgorithmPFSelection<mckl::BackendSEQ>());
    sampler.resample(mckl::Stratified);
    sampler.resample_threshold(0.5);
    const std::size_t n = std::min(tmax, sampler.particle().state().n());
    sampler.iterate(n);
    *T = n;
    sampler.read_ess_history(ess);
    std::vector<std::size_t> sz(n);
    sampler.read_size_history(sz.data());
    for (std::size_t i = 0; i != n; ++i)
        size[i] = sz[i];
    for (std::size_t j = 0; j != 4; ++j) {
        const double *c = sampler.particle().state().col_data(j);
        for (std::size_t i = 0; i != N; ++i)
            state[j * N + i] = c[i];
    }
    const double *w = sampler.particle().weight().data();
    for (std::size_t i = 0; i != N; ++i)
        weight[i] = w[i];
    return 0;
}

} // extern "C"
