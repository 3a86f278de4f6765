// host_api_test.cpp -- exercises the C++ host API (mckl_hip.hpp) the way the
// reference's example tests exercise theirs: KAT comparison
// (example/random/include/random_rng.hpp:80-104), bulk vs scalar draws
// (:234-331), resampling invariants and a sampler run.
// Usage: host_api_test <path to random_philox4x32.data>
#include <mckl_hip.hpp>

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <vector>

static int failures = 0;

#define CHECK(cond, what)                                                     \
    do {                                                                      \
        std::printf("%-48s %s\n", what, (cond) ? "Passed" : "Failed");        \
        if (!(cond))                                                          \
            ++failures;                                                       \
    } while (0)

template <typename T>
struct Dev {
    T *p = nullptr;
    explicit Dev(std::size_t n) { (void)hipMalloc(&p, n * sizeof(T)); }
    ~Dev() { (void)hipFree(p); }
    std::vector<T> host(std::size_t n) const
    {
        std::vector<T> h(n);
        (void)hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost);
        return h;
    }
};

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s random_philox4x32.data\n", argv[0]);
        return 2;
    }
    // 1. KAT of a default-constructed engine
    std::vector<std::uint32_t> kat;
    {
        std::ifstream f(argv[1]);
        std::uint32_t x;
        while (f >> x)
            kat.push_back(x);
    }
    {
        mckl::Philox4x32 rng;
        Dev<std::uint32_t> r(kat.size());
        mckl::rand(rng, kat.size(), r.p);
        CHECK(!kat.empty() && r.host(kat.size()) == kat, "Philox4x32 KAT");
        // seed(key) round trip (random_rng.hpp:97-103)
        mckl::Philox4x32::key_type k = {{rng(), rng()}};
        rng.seed(k);
        CHECK(rng.key() == k, "Philox4x32 seed(key) round trip");
    }
    // 2. scalar draws equal the bulk stream, across partial blocks
    {
        mckl::Threefry4x64 a(7), b(7);
        Dev<std::uint32_t> r(37);
        mckl::rand(a, 3, r.p);
        mckl::rand(a, 34, r.p + 3);
        std::vector<std::uint32_t> bulk = r.host(37), scal(37);
        for (auto &v : scal)
            v = b();
        CHECK(bulk == scal && a == b, "Threefry4x64 scalar == bulk");
        mckl::ARS c(3), d(3);
        c.discard(1001);
        for (int i = 0; i < 1001; ++i)
            d();
        CHECK(c() == d(), "ARS discard == scalar draws");
        std::stringstream ss;
        ss << c;
        mckl::ARS e(77);
        ss >> e;
        CHECK(!ss.fail() && e == c && e() == c(),
            "ARS operator<< / operator>> round trip");
        // the reference's text for Philox4x32 key {7, 9} after rand(5),
        // one scalar draw and discard(3) (tests/golden/engine_text.json)
        mckl::Philox4x32 p(std::array<std::uint32_t, 2>{{7, 9}});
        Dev<std::uint32_t> five(5);
        mckl::rand(p, 5, five.p);
        (void)p();
        p.discard(3);
        std::stringstream ps;
        ps << p;
        CHECK(ps.str() == "4 3170914008 3725014840 3378673226 1820808664 2 3 0 "
                          "2 7 9 1",
            "Philox4x32 operator<< matches the reference text");
    }
    // 3. Normal<double> moments
    {
        mckl::ARS rng(11);
        mckl::NormalDistribution<double> normal(2.5, 0.3);
        const std::size_t n = 1 << 20;
        Dev<double> r(n);
        mckl::rand(rng, normal, n, r.p);
        std::vector<double> h = r.host(n);
        double s = 0, s2 = 0;
        for (double v : h) {
            s += v;
            s2 += v * v;
        }
        const double mu = s / n, var = s2 / n - mu * mu;
        CHECK(std::fabs(mu - 2.5) < 5 * 0.3 / std::sqrt(double(n)) &&
                std::fabs(var - 0.09) < 0.001,
            "NormalDistribution<double>(2.5, 0.3) moments");
    }
    // 4. Weight and systematic resampling of equal weights
    {
        const std::size_t n = 5000;
        Dev<double> w(n), st(2), v(n);
        std::vector<double> zeros(n, 0.0);
        (void)hipMemcpy(v.p, zeros.data(), n * sizeof(double),
            hipMemcpyHostToDevice);
        mckl::Weight weight(n, w.p, st.p);
        weight.set_log(v.p);
        std::vector<double> s = st.host(2);
        CHECK(std::fabs(s[0] - double(n)) < 1e-9, "Weight::set_log ESS");
        CHECK(weight.ess() == s[0], "Weight::ess");
        mckl::Philox4x32 rng(5);
        Dev<std::uint32_t> rep(n), idx(n);
        mckl::ResampleSystematic()(n, n, rng, w.p, rep.p, idx.p);
        std::vector<std::uint32_t> r = rep.host(n), x = idx.host(n);
        bool ok = true;
        for (std::size_t i = 0; i != n; ++i)
            ok = ok && r[i] == 1 && x[i] == i;
        CHECK(ok, "ResampleSystematic on equal weights");
    }
    // 5. errors surface as exceptions
    {
        bool thrown = false;
        try {
            mckl::Philox4x32 rng;
            mckl::NormalDistribution<float> bad(0.0f, -1.0f);
            Dev<float> r(4);
            mckl::rand(rng, bad, 4, r.p);
        } catch (const mckl::RuntimeAssert &) {
            thrown = true;
        }
        CHECK(thrown, "invalid stddev raises RuntimeAssert");
    }
    // 5b. Particle::select checks the row numbers before the gather
    {
        mckl::Particle<double> p(8);
        std::vector<std::uint32_t> h = {0, 1, 2, 3, 4, 5, 6, 8};
        Dev<std::uint32_t> idx(8);
        (void)hipMemcpy(idx.p, h.data(), 8 * sizeof(std::uint32_t),
            hipMemcpyHostToDevice);
        bool thrown = false;
        try {
            p.select(8, idx.p);
        } catch (const mckl::RuntimeAssert &) {
            thrown = true;
        }
        h[7] = 7;
        (void)hipMemcpy(idx.p, h.data(), 8 * sizeof(std::uint32_t),
            hipMemcpyHostToDevice);
        p.select(8, idx.p);
        CHECK(thrown && p.size() == 8,
            "Particle::select rejects a row number >= size()");
    }
    // 6. plain sampler on the linear-Gaussian model
    {
        std::vector<double> y(30);
        for (std::size_t t = 0; t != y.size(); ++t)
            y[t] = std::sin(0.3 * t);
        mckl::SMCSamplerHIP sampler(MCKL_HIP_MODEL_LG, {{0.9, 1.0, 0.5}}, y,
            1 << 14, 3);
        sampler.resample_threshold(0.5);
        sampler.resample(mckl::Systematic);
        sampler.iterate();
        const double ll = sampler.log_likelihood();
        CHECK(std::isfinite(ll) && ll < 0, "SMCSamplerHIP log likelihood");
    }
    // 7. ResampleIndex trace-back equals the direct copy of the genealogy
    //    (example/algorithm/include/algorithm_resample_index.hpp:37-70,
    //    :108-135), both layouts
    {
        const std::size_t N = 1000, dim = 70;
        mckl::Philox4x32 rng(9);
        std::vector<std::vector<std::uint32_t>> index(dim);
        std::vector<std::uint32_t> draws(N * dim);
        {
            Dev<std::uint32_t> d(N * dim);
            mckl::rand(rng, N * dim, d.p);
            draws = d.host(N * dim);
        }
        mckl::ResampleIndex ri;
        Dev<std::uint32_t> di(N);
        for (std::size_t d = 0; d != dim; ++d) {
            index[d].resize(N);
            for (std::size_t i = 0; i != N; ++i)
                index[d][i] = draws[d * N + i] % N;
            (void)hipMemcpy(di.p, index[d].data(), N * 4,
                hipMemcpyHostToDevice);
            ri.push_back(N, di.p);
        }
        // direct copy: column d holds the value d*N + i, then every column
        // up to d is gathered through index[d]
        std::vector<std::uint32_t> state(N * dim), tmp(N * dim);
        for (std::size_t d = 0; d != dim; ++d) {
            for (std::size_t i = 0; i != N; ++i)
                state[i * dim + d] = static_cast<std::uint32_t>(d * N + i);
            tmp = state;
            for (std::size_t i = 0; i != N; ++i)
                for (std::size_t j = 0; j <= d; ++j)
                    state[i * dim + j] = tmp[index[d][i] * dim + j];
        }
        bool ok = ri.num_iter() == dim && ri.index_matrix_nrow() == N &&
            ri.index_matrix_ncol() == dim;
        for (int pass = 0; pass != 2; ++pass) {
            const auto layout = pass == 0 ? mckl::RowMajor : mckl::ColMajor;
            Dev<std::uint32_t> m(N * dim);
            std::uint32_t *end = ri.read_index_matrix(layout, m.p);
            std::vector<std::uint32_t> h = m.host(N * dim);
            ok = ok && end == m.p + N * dim;
            for (std::size_t i = 0; i != N; ++i)
                for (std::size_t j = 0; j != dim; ++j) {
                    const std::uint32_t e =
                        pass == 0 ? h[i * dim + j] : h[j * N + i];
                    ok = ok && state[i * dim + j] == j * N + e;
                }
        }
        ok = ok && ri.index(17) == state[17 * dim] % N;
        CHECK(ok, "ResampleIndex trace == direct copy");
    }
    // 8. SMCEstimator weighted sums and the sampler's moment estimator
    {
        const std::size_t n = 3000, d = 2;
        std::vector<double> u(n * d), w(n, 1.0 / n);
        for (std::size_t i = 0; i != n; ++i) {
            u[i * d] = double(i);
            u[i * d + 1] = 1.0;
        }
        Dev<double> du(n * d), dw(n), dr(d);
        (void)hipMemcpy(du.p, u.data(), n * d * 8, hipMemcpyHostToDevice);
        (void)hipMemcpy(dw.p, w.data(), n * 8, hipMemcpyHostToDevice);
        mckl::estimate(n, d, mckl::RowMajor, du.p, dw.p, dr.p);
        std::vector<double> r = dr.host(d);
        // sum_i i / n = (n - 1) / 2; sum_i 1 / n = 1 up to the rounding of
        // the weights 1/n
        CHECK(std::fabs(r[0] - 1499.5) < 1e-9 && std::fabs(r[1] - 1) < 1e-12,
            "estimate (weighted sums)");
        std::vector<double> y(20);
        for (std::size_t t = 0; t != y.size(); ++t)
            y[t] = std::cos(0.2 * t);
        mckl::SMCSamplerHIP s(MCKL_HIP_MODEL_LG, {{0.9, 1.0, 0.5}}, y, 4096, 5);
        s.estimate_moments(2);
        s.iterate();
        std::vector<double> e = s.estimates();
        bool ok = e.size() == 40;
        for (std::size_t t = 0; ok && t != 20; ++t)
            ok = e[2 * t + 1] >= e[2 * t] * e[2 * t];
        CHECK(ok, "SMCSamplerHIP moment estimates (E x^2 >= (E x)^2)");
    }
    // 9. PMMH with the device log-likelihood
    {
        std::vector<double> y(20);
        for (std::size_t t = 0; t != y.size(); ++t)
            y[t] = std::sin(0.3 * t);
        auto prior = [](const std::array<double, 3> &p) {
            return std::fabs(p[0]) < 1 && p[1] > 0 && p[2] > 0 ? 0.0 :
                                                                  -HUGE_VAL;
        };
        mckl::PMCMCMutationHIP mut(256, 20, prior, MCKL_HIP_MODEL_LG, y, 3);
        mut.update([](mckl::Philox4x32 &rng, std::array<double, 3> &p) {
            mckl::NormalDistribution<double> normal(0, 0.05);
            Dev<double> z(3);
            mckl::rand(rng, normal, 3, z.p);
            std::vector<double> h = z.host(3);
            for (int k = 0; k != 3; ++k)
                p[k] += h[k];
            return 0.0;
        });
        std::array<double, 3> param{{0.8, 1.0, 0.5}};
        std::size_t acc = 0;
        for (std::size_t it = 0; it != 30; ++it)
            acc += mut(it, param);
        CHECK(acc > 0 && acc < 29 && std::isfinite(mut.log_nc()) &&
                std::fabs(param[0]) < 1,
            "PMCMCMutationHIP chain");
    }
    // 10. SMCSampler<T> with evaluation objects equals the same steps done
    // by hand with the primitives (Weight, ResampleSystematic, select)
    {
        const std::size_t N = 5000, T = 9;
        mckl::NormalDistribution<double> normal(0, 1);
        mckl::Philox4x32 noise(77);
        Dev<double> v(N);
        mckl::SMCSampler<double> s(N, 1, mckl::Philox4x32(5));
        std::vector<int> order;
        s.selection([&](std::size_t t, mckl::Particle<double> &p) {
            if (t == 0)
                mckl::rand(noise, normal, N, p.state());
            mckl::rand(noise, normal, N, v.p);
            p.weight().add_log(v.p);
            order.push_back(0);
        });
        s.resample(mckl::Systematic);
        s.resample([&](std::size_t, mckl::Particle<double> &) {
            order.push_back(1);
        });
        s.mutation([&](std::size_t, mckl::Particle<double> &) {
            order.push_back(2);
        });
        s.selection_estimator(mckl::SMCEstimator<double>(1,
            [](std::size_t, std::size_t, mckl::Particle<double> &p,
                double *u) {
                (void)hipMemcpy(u, p.state(), p.size() * sizeof(double),
                    hipMemcpyDeviceToDevice);
            }));
        // log weights N(0, 1) per step: the ESS of one increment is about
        // N / e, of two about N / e^2, so every other step resamples
        s.resample_threshold(0.3);
        s.iterate(T);

        mckl::Philox4x32 noise2(77), rng2(5);
        Dev<double> v2(N), x(N), xb(N), w(N), st(2);
        Dev<std::uint32_t> rep(N), idx(N);
        Dev<double> r(1);
        mckl::Weight W(N, w.p, st.p);
        mckl::rand(noise2, normal, N, x.p);
        std::vector<double> ess, est;
        std::vector<int> want;
        for (std::size_t t = 0; t != T; ++t) {
            mckl::rand(noise2, normal, N, v2.p);
            W.add_log(v2.p);
            want.push_back(0);
            mckl::estimate(N, 1, mckl::MatrixLayout::RowMajor, x.p, w.p, r.p);
            est.push_back(r.host(1)[0]);
            ess.push_back(W.ess());
            if (ess.back() < 0.3 * N) {
                mckl::ResampleSystematic()(N, N, rng2, w.p, rep.p, idx.p);
                mckl::select_rows(N, 1, idx.p, x.p, xb.p);
                std::swap(x.p, xb.p);
                W.set_equal();
                want.push_back(1);
            }
            want.push_back(2);
        }
        std::vector<double> gess(T);
        s.read_ess_history(gess.begin());
        std::vector<double> gx(N);
        (void)hipMemcpy(gx.data(), s.particle().state(), N * sizeof(double),
            hipMemcpyDeviceToHost);
        const std::vector<double> sm = s.summary();
        bool same_est = sm.size() == T;
        for (std::size_t t = 0; same_est && t != T; ++t)
            same_est = sm[t] == est[t];
        std::size_t nres = 0;
        for (int o : want)
            nres += o == 1;
        CHECK(gess == ess && s.num_iter() == T,
            "SMCSampler<double> ESS history == the primitives'");
        CHECK(order == want && nres > 0 && nres < T,
            "SMCSampler<double> evaluation order");
        CHECK(gx == x.host(N), "SMCSampler<double> state == the primitives'");
        CHECK(same_est, "SMCSampler<double> estimates == the primitives'");
    }
    // 11. PMCMCMutation<Param, T> over SMCSampler<T>: the log normalising
    // constant is the parameter's own function here (a fixed log weight
    // increment per step), so it can be checked exactly, and a rejected step
    // restores the previous parameter and constant
    {
        using Param = std::array<double, 1>;
        const std::size_t N = 1000, M = 4;
        auto prior = [](const Param &p) {
            return p[0] > -3 && p[0] < 3 ? 0.0 : -HUGE_VAL;
        };
        mckl::PMCMCMutation<Param, double> mut(
            N, M, prior, 1, mckl::Philox4x32(13));
        std::vector<double> hv(N);
        Dev<double> dv(N);
        // log likelihood -p^2 / 2 per step in total: each step adds
        // log(sum W exp(v)) with v constant = -p^2 / (2 M)
        mut.pf().selection([&](std::size_t, mckl::Particle<double> &p) {
            const double c = -mut.state().param()[0] *
                mut.state().param()[0] / (2.0 * M);
            std::fill(hv.begin(), hv.end(), c);
            (void)hipMemcpy(dv.p, hv.data(), N * sizeof(double),
                hipMemcpyHostToDevice);
            p.weight().add_log(dv.p);
            mut.state().add_log_nc(c);
        });
        mut.pf().resample(mckl::Stratified);
        mut.pf().resample_threshold(0.5);
        mut.update([](mckl::Philox4x32 &rng, Param &p) {
            mckl::NormalDistribution<double> normal(0, 0.7);
            Dev<double> z(1);
            mckl::rand(rng, normal, 1, z.p);
            p[0] += z.host(1)[0];
            return 0.0;
        });
        Param param{{1.5}};
        std::size_t acc = 0;
        bool ok = true;
        for (std::size_t it = 0; it != 200; ++it) {
            const Param before = param;
            const std::size_t a = mut(it, param);
            acc += a;
            const double want = -param[0] * param[0] / 2.0;
            ok = ok && std::fabs(mut.state().log_nc() - want) < 1e-12 &&
                (a == 1 || param == before);
        }
        CHECK(ok && acc > 20 && acc < 190,
            "PMCMCMutation<Param, double> over SMCSampler<double>");
    }
    // 12. SMCBatchHIP equals SMCSamplerHIP filter by filter; SMCGradientHIP's
    // gradient matches central differences of its replayed function
    {
        std::vector<double> y(25);
        for (std::size_t t = 0; t != y.size(); ++t)
            y[t] = 0.8 * std::sin(0.4 * t);
        const std::vector<std::uint32_t> N{3000, 2049, 517};
        std::vector<double> th{0.9, 1.0, 0.5, 0.8, 0.7, 0.6, 0.95, 1.2, 0.4};
        std::vector<double> yy;
        for (int f = 0; f != 3; ++f)
            yy.insert(yy.end(), y.begin(), y.end());
        const std::vector<std::uint64_t> seed{11, 12, 13};
        mckl::SMCBatchHIP b(MCKL_HIP_MODEL_LG, N, th, y.size(), yy, seed, 0.5,
            MCKL_HIP_SYSTEMATIC);
        b.run();
        const std::vector<double> lb = b.log_likelihood();
        bool same = lb.size() == 3;
        for (int f = 0; f != 3 && same; ++f) {
            mckl::SMCSamplerHIP s(MCKL_HIP_MODEL_LG,
                {{th[3 * f], th[3 * f + 1], th[3 * f + 2]}}, y, N[f], seed[f]);
            s.resample_threshold(0.5);
            s.resample(mckl::Systematic);
            s.iterate();
            same = s.log_likelihood() == lb[f];
        }
        CHECK(same, "SMCBatchHIP == SMCSamplerHIP per filter");
        double tms = 0.0;
        const mckl_hip_tune tt = b.autotune(std::string(), &tms);
        b.run();
        CHECK(b.log_likelihood() == lb && tms > 0 && tt.block >= 256,
            "SMCBatchHIP::autotune keeps the bits");

        const std::vector<double> th2{0.9, 1.0, 0.5, 0.85, 1.1, 0.45};
        mckl::SMCGradientHIP g(MCKL_HIP_MODEL_LG, 2, 2000, y.size(), y, false,
            {3, 4}, 0.5, MCKL_HIP_STRATIFIED, 7, true);
        std::vector<double> ll, gr;
        g.run(th2, ll, gr);
        const std::vector<double> v = g.replay(th2);
        bool ok = v == ll;
        const double h = 1e-5;
        for (int j = 0; j != 3 && ok; ++j) {
            std::vector<double> tp(th2), tm(th2);
            for (int p = 0; p != 2; ++p) {
                tp[3 * p + j] += h;
                tm[3 * p + j] -= h;
            }
            const std::vector<double> fp = g.replay(tp), fm = g.replay(tm);
            for (int p = 0; p != 2; ++p) {
                const double fd = (fp[p] - fm[p]) / (2 * h);
                ok = ok && std::fabs(fd - gr[3 * p + j]) <=
                        1e-4 * std::max(1.0, std::fabs(gr[3 * p + j]));
            }
        }
        CHECK(ok, "SMCGradientHIP replay == forward and gradient == FD");
    }
    std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
    return failures ? 1 : 0;
}
