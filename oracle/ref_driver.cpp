// ref_driver.cpp -- thin C entry points over the reference's own headers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/_ref/ from
// the reference headers where they lie (header-only library; nothing is
// copied).  Used to validate the CPU restatement (mckl_oracle.c), to write
// the committed fixtures under tests/golden/, and as the "reference" CPU
// baseline in bench.py.  The product never links it.
//
// Pinned flags (oracle/Makefile): -std=c++14 -O2 -ffp-contract=off, no -march
// => SSE2 paths, no FMA in muladd (math/vmf.hpp:646-722), software AES for ARS
// (random/aes.hpp:76-104), MCKL_REQUIRE_ENDIANNESS_NEUTURAL=1 as the
// reference's examples are built (example/CMakeLists.txt:34).

#include <mckl/algorithm/resample.hpp>
#include <mckl/core/weight.hpp>
#include <mckl/random/aes.hpp>
#include <mckl/random/normal_distribution.hpp>
#include <mckl/random/uniform_bits_distribution.hpp>
#include <mckl/random/philox.hpp>
#include <mckl/random/threefry.hpp>
#include <mckl/random/u01_distribution.hpp>
#include <mckl/random/uniform_real_distribution.hpp>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace {

template <typename RNG>
void set_key(RNG &rng, const std::uint64_t *key)
{
    typename RNG::key_type k;
    for (std::size_t i = 0; i != k.size(); ++i)
        k[i] = static_cast<typename RNG::key_type::value_type>(key[i]);
    rng.seed(k);
}

// Engine state as (ctr words, index) parsed from the reference's own
// operator<< (random/counter.hpp:235-250): "result... ctr... gen... index".
template <typename RNG>
void get_state(const RNG &rng, std::uint64_t *ctr, unsigned *index)
{
    auto c = rng.ctr();
    for (std::size_t i = 0; i != c.size(); ++i)
        ctr[i] = c[i];
    std::ostringstream os;
    os << rng;
    std::string s = os.str();
    std::size_t p = s.find_last_of(' ');
    *index = static_cast<unsigned>(std::stoul(s.substr(p + 1)));
}

template <typename RNG, typename Fn>
int with_engine(int rbits, const std::uint64_t *key, Fn &&fn);

template <typename Fn>
int dispatch(int kind, int rbits, const std::uint64_t *key, Fn &&fn)
{
    switch (kind) {
        case 1:
            if (rbits == 32) {
                mckl::Philox4x32 rng;
                set_key(rng, key);
                return fn(rng);
            } else {
                mckl::Philox4x32_64 rng;
                set_key(rng, key);
                return fn(rng);
            }
        case 2:
            if (rbits == 32) {
                mckl::Threefry4x64 rng;
                set_key(rng, key);
                return fn(rng);
            } else {
                mckl::Threefry4x64_64 rng;
                set_key(rng, key);
                return fn(rng);
            }
        case 3:
            if (rbits == 32) {
                mckl::ARS rng;
                set_key(rng, key);
                return fn(rng);
            } else {
                mckl::ARS_64 rng;
                set_key(rng, key);
                return fn(rng);
            }
    }
    return -1;
}

} // namespace

extern "C" {

// A sequence of bulk rand calls (sizes[0..ncalls)), output concatenated.
// ops: size > 0 is rand(rng, size, r); size < 0 is discard(-size); size == 0
// is one scalar rng() call.  Final state returned in ctr/index.
int ref_engine_ops(int kind, int rbits, const std::uint64_t *key,
    const std::uint64_t *ctr0, int ncalls, const long long *sizes, void *out,
    std::uint64_t *ctr, unsigned *index)
{
    return dispatch(kind, rbits, key, [&](auto &rng) {
        using R = typename std::remove_reference_t<decltype(rng)>::result_type;
        if (ctr0) {
            auto c = rng.ctr();
            for (std::size_t i = 0; i != c.size(); ++i)
                c[i] = ctr0[i];
            rng.ctr(c);
        }
        R *r = static_cast<R *>(out);
        for (int i = 0; i != ncalls; ++i) {
            if (sizes[i] > 0) {
                mckl::rand(rng, static_cast<std::size_t>(sizes[i]), r);
                r += sizes[i];
            } else if (sizes[i] < 0) {
                rng.discard(static_cast<typename std::remove_reference_t<
                        decltype(rng)>::skip_type>(-sizes[i]));
            } else {
                *r++ = rng();
            }
        }
        get_state(rng, ctr, index);
        return 0;
    });
}

// The engine's textual state (operator<<, random/counter.hpp:235-249)
// after the same call sequence as ref_engine_ops; written NUL-terminated to
// `text` (capacity `cap`).  Returns the length, or -1 if it does not fit.
int ref_engine_text(int kind, int rbits, const std::uint64_t *key, int ncalls,
    const long long *sizes, char *text, std::size_t cap)
{
    return dispatch(kind, rbits, key, [&](auto &rng) {
        using R = typename std::remove_reference_t<decltype(rng)>::result_type;
        std::vector<R> buf(1);
        for (int i = 0; i != ncalls; ++i) {
            if (sizes[i] > 0) {
                buf.resize(static_cast<std::size_t>(sizes[i]));
                mckl::rand(rng, static_cast<std::size_t>(sizes[i]), buf.data());
            } else if (sizes[i] < 0) {
                rng.discard(static_cast<typename std::remove_reference_t<
                        decltype(rng)>::skip_type>(-sizes[i]));
            } else {
                (void)rng();
            }
        }
        std::ostringstream os;
        os << rng;
        const std::string s = os.str();
        if (s.size() + 1 > cap)
            return -1;
        std::memcpy(text, s.c_str(), s.size() + 1);
        return static_cast<int>(s.size());
    });
}

// pre raw result words, then rand(rng, U01{CC,CO,OC,OO,Canonical}<T>, n, r)
// (rbits 32 or 64: the engine's result width)
int ref_u01(int kind, int rbits, const std::uint64_t *key, int interval, int dbl,
    std::size_t pre, std::size_t n, void *out, std::uint64_t *ctr,
    unsigned *index)
{
    return dispatch(kind, rbits, key, [&](auto &rng) {
        using R = typename std::remove_reference_t<decltype(rng)>::result_type;
        std::vector<R> skip(pre + 1);
        mckl::rand(rng, pre, skip.data());
        auto run = [&](auto *r) {
            using T = std::remove_pointer_t<decltype(r)>;
            switch (interval) {
                case 0: {
                    mckl::U01CCDistribution<T> d;
                    mckl::rand(rng, d, n, r);
                } break;
                case 1: {
                    mckl::U01CODistribution<T> d;
                    mckl::rand(rng, d, n, r);
                } break;
                case 2: {
                    mckl::U01OCDistribution<T> d;
                    mckl::rand(rng, d, n, r);
                } break;
                case 3: {
                    mckl::U01OODistribution<T> d;
                    mckl::rand(rng, d, n, r);
                } break;
                default: {
                    mckl::U01CanonicalDistribution<T> d;
                    mckl::rand(rng, d, n, r);
                } break;
            }
        };
        if (dbl)
            run(static_cast<double *>(out));
        else
            run(static_cast<float *>(out));
        get_state(rng, ctr, index);
        return 0;
    });
}

// pre raw result words, then rand(rng, UniformBitsDistribution<U>, n, r)
// for U of ubits = 16, 32 or 64 bits
int ref_uniform_bits(int kind, int rbits, const std::uint64_t *key,
    int ubits, std::size_t pre, std::size_t n, void *out, std::uint64_t *ctr,
    unsigned *index)
{
    return dispatch(kind, rbits, key, [&](auto &rng) {
        using R = typename std::remove_reference_t<decltype(rng)>::result_type;
        std::vector<R> skip(pre + 1);
        mckl::rand(rng, pre, skip.data());
        auto run = [&](auto *r) {
            using U = std::remove_pointer_t<decltype(r)>;
            mckl::UniformBitsDistribution<U> d;
            mckl::rand(rng, d, n, r);
        };
        if (ubits == 16)
            run(static_cast<std::uint16_t *>(out));
        else if (ubits == 32)
            run(static_cast<std::uint32_t *>(out));
        else
            run(static_cast<std::uint64_t *>(out));
        get_state(rng, ctr, index);
        return 0;
    });
}

// pre raw result words, then rand(rng, NormalDistribution<T>(mean, sd), n, r)
int ref_normal(int kind, int rbits, const std::uint64_t *key, int dbl, std::size_t pre,
    std::size_t n, double mean, double sd, void *out, std::uint64_t *ctr,
    unsigned *index)
{
    return dispatch(kind, rbits, key, [&](auto &rng) {
        using R = typename std::remove_reference_t<decltype(rng)>::result_type;
        std::vector<R> skip(pre + 1);
        mckl::rand(rng, pre, skip.data());
        if (dbl) {
            mckl::NormalDistribution<double> d(mean, sd);
            mckl::rand(rng, d, n, static_cast<double *>(out));
        } else {
            mckl::NormalDistribution<float> d(
                static_cast<float>(mean), static_cast<float>(sd));
            mckl::rand(rng, d, n, static_cast<float *>(out));
        }
        get_state(rng, ctr, index);
        return 0;
    });
}

// pre raw result words, then rand(rng, UniformRealDistribution<T>(a, b), n,
// r)
int ref_uniform_real(int kind, int rbits, const std::uint64_t *key, int dbl,
    std::size_t pre, std::size_t n, double a, double b, void *out,
    std::uint64_t *ctr, unsigned *index)
{
    return dispatch(kind, rbits, key, [&](auto &rng) {
        using R = typename std::remove_reference_t<decltype(rng)>::result_type;
        std::vector<R> skip(pre + 1);
        mckl::rand(rng, pre, skip.data());
        if (dbl) {
            mckl::UniformRealDistribution<double> d(a, b);
            mckl::rand(rng, d, n, static_cast<double *>(out));
        } else {
            mckl::UniformRealDistribution<float> d(
                static_cast<float>(a), static_cast<float>(b));
            mckl::rand(rng, d, n, static_cast<float *>(out));
        }
        get_state(rng, ctr, index);
        return 0;
    });
}

// Weight::set_log (mode 0) or Weight::add_log (mode 1) on W (in/out).
int ref_weight(int mode, std::size_t n, const double *v, double *w, double *ess)
{
    mckl::Weight weight(n);
    if (mode == 1)
        weight.set_exact(*ess, w);
    if (mode == 0)
        weight.set_log(v);
    else
        weight.add_log(v);
    weight.read(w);
    *ess = weight.ess();
    return 0;
}

// U01 sequences: scheme 0 sorted (multinomial), 1 stratified, 2 systematic
// (ResampleScheme numbering)
int ref_u01_sequence(int scheme, int kind, const std::uint64_t *key,
    std::size_t n, double *r, std::uint64_t *ctr, unsigned *index)
{
    const int base = scheme >= 3 ? scheme - 3 : scheme; // residual family
    return dispatch(kind, 32, key, [&](auto &rng) {
        if (base == 0)
            mckl::u01_rand_sorted(rng, n, r);
        else if (base == 1)
            mckl::u01_rand_stratified(rng, n, r);
        else
            mckl::u01_rand_systematic(rng, n, r);
        get_state(rng, ctr, index);
        return 0;
    });
}

// the transforms of given standard uniforms into the schemes' sequences
// (u01_trans_sorted/stratified/systematic, algorithm/resample.hpp:141-212);
// scheme numbering as above, systematic reads u01[0] only
int ref_u01_trans(int scheme, std::size_t n, const double *u01, double *r)
{
    if (scheme == 0)
        mckl::u01_trans_sorted(n, u01, r);
    else if (scheme == 1)
        mckl::u01_trans_stratified(n, u01, r);
    else
        mckl::u01_trans_systematic(n, u01, r);
    return 0;
}

// resample_trans_residual (algorithm/resample.hpp:333-364): normalised
// residuals and integral parts of M W_i; returns R = M - sum of the parts
std::size_t ref_trans_residual(std::size_t N, std::size_t M, const double *w,
    double *resid, std::uint64_t *integ)
{
    std::vector<std::size_t> in(N);
    const std::size_t R =
        mckl::resample_trans_residual(N, M, w, resid, in.data());
    for (std::size_t i = 0; i != N; ++i)
        integ[i] = in[i];
    return R;
}

int ref_resample(int scheme, std::size_t N, std::size_t M, int kind,
    const std::uint64_t *key, const double *w, std::uint64_t *rep,
    std::uint64_t *idx)
{
    return dispatch(kind, 32, key, [&](auto &rng) {
        std::vector<std::size_t> r(N + 1), x(M + 1);
        switch (scheme) {
            case 0:
                mckl::ResampleMultinomial()(N, M, rng, w, r.data());
                break;
            case 1:
                mckl::ResampleStratified()(N, M, rng, w, r.data());
                break;
            case 2:
                mckl::ResampleSystematic()(N, M, rng, w, r.data());
                break;
            case 3:
                mckl::ResampleResidual()(N, M, rng, w, r.data());
                break;
            case 4:
                mckl::ResampleResidualStratified()(N, M, rng, w, r.data());
                break;
            default:
                mckl::ResampleResidualSystematic()(N, M, rng, w, r.data());
        }
        mckl::resample_trans_rep_index(N, M, r.data(), x.data());
        for (std::size_t i = 0; i != N; ++i)
            rep[i] = r[i];
        for (std::size_t i = 0; i != M; ++i)
            idx[i] = x[i];
        return 0;
    });
}

int ref_trans_u01_rep(std::size_t N, std::size_t M, const double *w,
    const double *u, std::uint64_t *rep)
{
    std::vector<std::size_t> r(N + 1);
    mckl::resample_trans_u01_rep(N, M, w, u, r.data());
    for (std::size_t i = 0; i != N; ++i)
        rep[i] = r[i];
    return 0;
}

int ref_rep_index(
    std::size_t N, std::size_t M, const std::uint64_t *rep, std::uint64_t *idx)
{
    std::vector<std::size_t> r(rep, rep + N), x(M + 1);
    mckl::resample_trans_rep_index(N, M, r.data(), x.data());
    for (std::size_t i = 0; i != M; ++i)
        idx[i] = x[i];
    return 0;
}

// ResampleIndex<std::uint32_t> (algorithm/resample.hpp:597-832) driven op
// by op; callers pass only arguments the reference's runtime_asserts accept.
static mckl::ResampleIndex<std::uint32_t> *g_ri = nullptr;

int ref_ri_new()
{
    delete g_ri;
    g_ri = new mckl::ResampleIndex<std::uint32_t>();
    return 0;
}

int ref_ri_push(std::size_t N, const std::uint32_t *idx)
{
    if (idx)
        g_ri->push_back(N, idx);
    else
        g_ri->push_back(N);
    return 0;
}

int ref_ri_reset()
{
    g_ri->reset();
    return 0;
}

int ref_ri_clear()
{
    g_ri->clear();
    return 0;
}

// which = -2: num_iter(); -1: size(); 0: size(iter)
std::size_t ref_ri_size(int which, std::size_t iter)
{
    if (which == -2)
        return g_ri->num_iter();
    return which == -1 ? g_ri->size() : g_ri->size(iter);
}

std::size_t ref_ri_nrow(std::size_t iter_back)
{
    return g_ri->index_matrix_nrow(iter_back);
}

int ref_ri_index(std::size_t n, const std::uint32_t *id,
    std::size_t iter_back, std::size_t iter, std::uint32_t *out)
{
    for (std::size_t j = 0; j != n; ++j)
        out[j] = g_ri->index(id[j], iter_back, iter);
    return 0;
}

int ref_ri_matrix(
    int layout, std::size_t iter_back, std::size_t iter, std::uint32_t *out)
{
    g_ri->read_index_matrix(
        layout == 101 ? mckl::RowMajor : mckl::ColMajor, out, iter_back, iter);
    return 0;
}

// CPU baseline: time reps of `rand(Philox4x32, n, u32*)` followed by
// rand(U01OO<double>, n) on the same stream (fresh engine each rep), single
// thread.  Returns seconds per rep (median).
double ref_time_philox_u01(const std::uint64_t *key, std::size_t n, int reps,
    std::uint32_t *u32, double *f64)
{
    std::vector<double> t;
    for (int k = 0; k < reps; ++k) {
        mckl::Philox4x32 rng;
        set_key(rng, key);
        auto t0 = std::chrono::steady_clock::now();
        mckl::rand(rng, n, u32);
        mckl::Philox4x32 rng2;
        set_key(rng2, key);
        mckl::U01OODistribution<double> d;
        mckl::rand(rng2, d, n, f64);
        auto t1 = std::chrono::steady_clock::now();
        t.push_back(std::chrono::duration<double>(t1 - t0).count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

// The same step split over nthreads threads.  Thread i fills elements
// [lo, hi) of both outputs from its own engines, each moved to element lo by
// discard (one 32-bit result word per value: U01 doubles from a 32-bit engine
// take one word each, u01_distribution.hpp:95-113); lo is a multiple of
// 8, so every thread starts on a counter-block boundary and the outputs equal
// the single-thread ones.  Returns seconds per rep (median), thread start-up
// included.
double ref_time_philox_u01_mt(const std::uint64_t *key, std::size_t n,
    int reps, int nthreads, std::uint32_t *u32, double *f64)
{
    const std::size_t nt = static_cast<std::size_t>(nthreads < 1 ? 1 : nthreads);
    const std::size_t chunk = ((n + nt - 1) / nt + 7) / 8 * 8;
    std::vector<double> t;
    for (int k = 0; k < reps; ++k) {
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> th;
        for (std::size_t i = 0; i < nt; ++i) {
            const std::size_t lo = std::min(n, i * chunk);
            const std::size_t hi = std::min(n, lo + chunk);
            th.emplace_back([=]() {
                mckl::Philox4x32 a;
                set_key(a, key);
                a.discard(lo);
                mckl::rand(a, hi - lo, u32 + lo);
                mckl::Philox4x32 b;
                set_key(b, key);
                b.discard(lo);
                mckl::U01OODistribution<double> d;
                mckl::rand(b, d, hi - lo, f64 + lo);
            });
        }
        for (auto &x : th)
            x.join();
        auto t1 = std::chrono::steady_clock::now();
        t.push_back(std::chrono::duration<double>(t1 - t0).count());
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

// Independent Normal streams on nthreads threads: stream s (key {s + 1, 0,
// 0, 0}, 32-bit results) fills out[s * n, (s + 1) * n) with rand(rng,
// NormalDistribution<T>(mean, sd), n, r); thread i takes streams i, i +
// nthreads, ...  Returns wall seconds for the whole pass, or -1 on a bad
// engine kind.
double ref_time_normal_streams(int kind, int dbl, int nstreams, std::size_t n,
    double mean, double sd, int nthreads, void *out)
{
    const int nt = nthreads < 1 ? 1 : nthreads;
    std::vector<int> rc(static_cast<std::size_t>(nt), 0);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) {
        th.emplace_back([=, &rc]() {
            for (int s = i; s < nstreams; s += nt) {
                const std::uint64_t key[4] = {
                    static_cast<std::uint64_t>(s + 1), 0, 0, 0};
                const std::size_t off = static_cast<std::size_t>(s) * n;
                rc[static_cast<std::size_t>(i)] |=
                    dispatch(kind, 32, key, [&](auto &rng) {
                        if (dbl) {
                            mckl::NormalDistribution<double> d(mean, sd);
                            mckl::rand(rng, d, n,
                                static_cast<double *>(out) + off);
                        } else {
                            mckl::NormalDistribution<float> d(
                                static_cast<float>(mean),
                                static_cast<float>(sd));
                            mckl::rand(rng, d, n,
                                static_cast<float *>(out) + off);
                        }
                        return 0;
                    });
            }
        });
    }
    for (auto &x : th)
        x.join();
    auto t1 = std::chrono::steady_clock::now();
    for (int r : rc)
        if (r != 0)
            return -1.0;
    return std::chrono::duration<double>(t1 - t0).count();
}

} // extern "C"
