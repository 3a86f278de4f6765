// algorithm_pf.hip -- the reference's particle-filter example
// (example/algorithm/include/algorithm_pf.hpp:38-150, BackendSEQ) written
// against the device API, for tests/test_gpu_pf.py.  Test program.
//
// State: SMCSampler<double> with 4 ColMajor columns p, q, s, t.  Selection
// at iteration t: Normal fills on the range's engine (BackendSEQ runs one
// range of all N particles, whose engine is the RNG set's engine 0), in the
// example's order, then one device functor for the arithmetic of its vmf
// calls, in their order and with their roundings (muladd as a * b + c, two
// roundings, as the reference without FMA), and Weight::add_log.
// Resampling: Stratified at threshold 0.5 on Particle::rng().  Engines are
// keyed as the reference's SeedGenerator keys them with internal seed
// `seed`: the N engines of the RNG set, then Particle::rng().
//
// usage: algorithm_pf <data> <N> <seed> <tmax> <out> [time]
// out: T, then ess[T], size[T], state[4 N] (column-major), weight[N], as
// little-endian doubles; then keys[4] (RNG set engine 0, Particle::rng())
// as u64.  With `time`, the example's timed second run follows (clear(),
// then iterate again, algorithm_pf.hpp:192-197) and its wall time is
// printed as "iterate_seconds <s>".

#include <mckl_hip_eval.hpp>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

namespace {

constexpr double kScale = 10;
constexpr double kNuInv = 0.1;
constexpr double kCoeff = -5.5;
constexpr double kDelta = 0.1;

// iterations after the first: the increments a1..a4 were drawn in the
// example's order (rp, rv, rp, rv); the columns are updated as its muladd
// and add calls update them
struct PFMove {
    double *p, *q, *s, *t;
    const double *a1, *a2, *a3, *a4;

    __device__ void operator()(std::size_t i) const
    {
        double w = __dadd_rn(__dmul_rn(kDelta, s[i]), a1[i]); // muladd
        p[i] = __dadd_rn(w, p[i]);                               // add
        s[i] = __dadd_rn(a2[i], s[i]);
        w = __dadd_rn(__dmul_rn(kDelta, t[i]), a3[i]);
        q[i] = __dadd_rn(w, q[i]);
        t[i] = __dadd_rn(a4[i], t[i]);
    }
};

// the log-weight increment: sub, mul, sqr, mul, log1p on each coordinate,
// then add and mul
struct PFWeight {
    const double *p, *q;
    double x, y;
    double *w;
    mckl::DeviceMath m;

    __device__ void operator()(std::size_t i) const
    {
        double a = __dsub_rn(p[i], x);
        a = __dmul_rn(kScale, a);
        a = __dmul_rn(a, a);
        a = __dmul_rn(a, kNuInv);
        a = m.log1p(a);
        double b = __dsub_rn(q[i], y);
        b = __dmul_rn(kScale, b);
        b = __dmul_rn(b, b);
        b = __dmul_rn(b, kNuInv);
        b = m.log1p(b);
        w[i] = __dmul_rn(kCoeff, __dadd_rn(a, b));
    }
};

mckl::Philox4x32 engine_of(const std::vector<std::uint64_t> &keys, std::size_t e)
{
    mckl::Philox4x32::key_type k;
    k[0] = static_cast<std::uint32_t>(keys[4 * e]);
    k[1] = static_cast<std::uint32_t>(keys[4 * e + 1]);
    return mckl::Philox4x32(k);
}

} // namespace

int main(int argc, char **argv)
{
    if (argc != 6 && argc != 7) {
        std::fprintf(stderr, "usage: %s data N seed tmax out [time]\n",
            argv[0]);
        return 2;
    }
    std::vector<double> xs, ys;
    {
        std::ifstream data(argv[1]);
        double x = 0, y = 0;
        while (data >> x >> y) {
            xs.push_back(x);
            ys.push_back(y);
        }
    }
    const std::size_t N = std::strtoull(argv[2], nullptr, 10);
    const std::uint64_t seed = std::strtoull(argv[3], nullptr, 10);
    const std::size_t T =
        std::min<std::size_t>(std::strtoull(argv[4], nullptr, 10), xs.size());
    if (N == 0 || T == 0) {
        std::fprintf(stderr, "algorithm_pf: empty run\n");
        return 2;
    }
    try {
        // the SeedGenerator's keys: RNG set engines 0..N-1, then rng()
        mckl::DeviceArray<std::uint64_t> dkeys(4 * (N + 1));
        mckl::hip_check(mckl_hip_seed_keys(MCKL_HIP_PHILOX4X32, seed, 1, 0, 1,
            N + 1, dkeys.data(), mckl::current_stream()));
        std::vector<std::uint64_t> keys(4 * (N + 1));
        if (hipMemcpy(keys.data(), dkeys.data(), keys.size() * 8,
                hipMemcpyDeviceToHost) != hipSuccess)
            throw mckl::RuntimeAssert("reading the keys failed");
        mckl::Philox4x32 range_rng = engine_of(keys, 0);

        mckl::SMCSampler<double> sampler(
            N, 4, engine_of(keys, N), mckl::ColMajor);
        mckl::DeviceArray<double> w(N), a1(N), a2(N), a3(N), a4(N);
        const mckl::DeviceMath dm = mckl::device_math();

        sampler.selection([&](std::size_t iter,
                              mckl::Particle<double> &particle) {
            double *p = particle.col_data(0);
            double *q = particle.col_data(1);
            double *s = particle.col_data(2);
            double *t = particle.col_data(3);
            if (iter == 0) {
                mckl::NormalDistribution<double> rp(0, 2);
                mckl::NormalDistribution<double> rv(0, 1);
                mckl::rand(range_rng, rp, N, p);
                mckl::rand(range_rng, rp, N, q);
                mckl::rand(range_rng, rv, N, s);
                mckl::rand(range_rng, rv, N, t);
            } else {
                mckl::NormalDistribution<double> rp(0, std::sqrt(0.02));
                mckl::NormalDistribution<double> rv(0, std::sqrt(0.001));
                mckl::rand(range_rng, rp, N, a1.data());
                mckl::rand(range_rng, rv, N, a2.data());
                mckl::rand(range_rng, rp, N, a3.data());
                mckl::rand(range_rng, rv, N, a4.data());
                mckl::eval_each(N, PFMove{p, q, s, t, a1.data(), a2.data(),
                                       a3.data(), a4.data()});
            }
            mckl::eval_each(
                N, PFWeight{p, q, xs[iter], ys[iter], w.data(), dm});
            particle.weight().add_log(w.data());
        });
        sampler.resample(mckl::Stratified);
        sampler.resample_threshold(0.5);
        sampler.iterate(T);

        std::vector<double> ess(T);
        std::vector<std::size_t> sz(T);
        sampler.read_ess_history(ess.begin());
        sampler.read_size_history(sz.begin());
        std::vector<double> state(4 * N), weight(N);
        if (hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(state.data(), sampler.particle().state(), 4 * N * 8,
                hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(weight.data(), sampler.particle().weight().data(), N * 8,
                hipMemcpyDeviceToHost) != hipSuccess)
            throw mckl::RuntimeAssert("reading the results failed");

        std::FILE *f = std::fopen(argv[5], "wb");
        if (!f)
            throw mckl::RuntimeAssert("cannot open the output file");
        const double td = static_cast<double>(T);
        std::fwrite(&td, 8, 1, f);
        std::fwrite(ess.data(), 8, T, f);
        for (std::size_t v : sz) {
            const double d = static_cast<double>(v);
            std::fwrite(&d, 8, 1, f);
        }
        std::fwrite(state.data(), 8, 4 * N, f);
        std::fwrite(weight.data(), 8, N, f);
        const std::uint64_t k[4] = {
            keys[0], keys[1], keys[4 * N], keys[4 * N + 1]};
        std::fwrite(k, 8, 4, f);
        std::fclose(f);

        if (argc == 7) {
            sampler.clear();
            const auto t0 = std::chrono::steady_clock::now();
            sampler.iterate(T);
            if (hipDeviceSynchronize() != hipSuccess)
                throw mckl::RuntimeAssert("the timed run failed");
            const auto t1 = std::chrono::steady_clock::now();
            std::printf("iterate_seconds %.6f\n",
                std::chrono::duration<double>(t1 - t0).count());
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "algorithm_pf: %s\n", e.what());
        return 1;
    }
    return 0;
}
