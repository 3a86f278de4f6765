// crmath.hpp -- correctly rounded double-precision log, log1p, sin and cos
// for the Normal fills and for device evaluation objects.
//
// The reference's NormalDistribution<double> takes std::log, std::sin and
// std::cos of each draw (random/normal_distribution.hpp:50-78,
// math/vmf.hpp:1024,1043-1077).  Those libm results are within one ulp but
// not always correctly rounded, and no two libms misround the same inputs;
// so the device computes the correctly rounded value of each function, and a
// result can differ from the reference only where the reference's libm
// misrounds.  Each function is evaluated in double-double arithmetic (FMA
// based, error below 2^-98 relative) and rounded once, which is the
// correctly rounded result unless the exact value lies within 2^-98 of a
// rounding boundary.
//
// Method (from the mathematics, not from any libm):
//   log x = e ln2 + log c_j + log1p(r),  x = 2^e f, f in [0.75, 1.5),
//           c_j = j/1024 nearest to f, r = (f - c_j)/c_j, |r| <= 2^-10.58,
//           log1p by its Taylor series (10 terms, the first five in
//           double-double), log c_j from a table built by the atanh series.
//   log1p x = the series for |x| < 2^-11, else log(s) + log1p(e/s) where
//           1 + x = s + e exactly.
//   sin/cos x, 0 <= x <= 8: x = k pi/2 + r (pi/2 in three 50-bit pieces,
//           so r is exact to 2^-150), r = a_j + s with a_j = j/64,
//           |s| <= 1/128, sin/cos s by Taylor series, the angle-addition
//           formulas with sin/cos a_j from a table built by the Taylor
//           series, and the quadrant from k.
// The functions are __host__ __device__ so the same code is checked on the
// CPU against quad-precision values.
#pragma once

#include <cstdint>
#include <cstring>

namespace mckl_hip {
namespace crm {

struct ddv {
    double hi, lo;
};

constexpr int kLogJ0 = 768;   // c_j = j / 1024, j = 768 .. 1536
constexpr int kLogJ1 = 1536;
constexpr int kLogN = kLogJ1 - kLogJ0 + 1;
constexpr int kTrigN = 52;    // a_j = j / 64, j = 0 .. 51 (0.797 > pi/4)

// table of log c_j and sin/cos a_j, each a double-double (hi, lo)
struct Tables {
    double log_c[2 * kLogN];
    double sin_a[2 * kTrigN];
    double cos_a[2 * kTrigN];
};

// ---- double-double building blocks ----
__host__ __device__ inline ddv two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return ddv{s, (a - (s - bb)) + (b - bb)};
}

// |a| >= |b| (or a == 0)
__host__ __device__ inline ddv fast_two_sum(double a, double b)
{
    const double s = a + b;
    return ddv{s, b - (s - a)};
}

__host__ __device__ inline ddv two_prod(double a, double b)
{
    const double p = a * b;
    return ddv{p, __builtin_fma(a, b, -p)};
}

// a + b, no catastrophic cancellation between a and b
__host__ __device__ inline ddv add(ddv a, ddv b)
{
    const ddv s = two_sum(a.hi, b.hi);
    return fast_two_sum(s.hi, s.lo + (a.lo + b.lo));
}

// a + b with cancellation allowed (two error-free sums)
__host__ __device__ inline ddv add_acc(ddv a, ddv b)
{
    ddv s = two_sum(a.hi, b.hi);
    const ddv t = two_sum(a.lo, b.lo);
    s = fast_two_sum(s.hi, s.lo + t.hi);
    return fast_two_sum(s.hi, s.lo + t.lo);
}

__host__ __device__ inline ddv mul(ddv a, ddv b)
{
    ddv p = two_prod(a.hi, b.hi);
    p.lo = __builtin_fma(a.hi, b.lo, p.lo);
    p.lo = __builtin_fma(a.lo, b.hi, p.lo);
    return fast_two_sum(p.hi, p.lo);
}

__host__ __device__ inline ddv mul_d(ddv a, double b)
{
    ddv p = two_prod(a.hi, b);
    p.lo = __builtin_fma(a.lo, b, p.lo);
    return fast_two_sum(p.hi, p.lo);
}

__host__ __device__ inline ddv div_d(ddv a, double b)
{
    const double q = a.hi / b;
    // a.hi - q b exactly, plus a.lo
    const double r = __builtin_fma(-q, b, a.hi) + a.lo;
    return fast_two_sum(q, r / b);
}

__host__ __device__ inline ddv neg(ddv a) { return ddv{-a.hi, -a.lo}; }

__host__ __device__ inline uint64_t as_u64(double x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof(u));
    return u;
}

__host__ __device__ inline double as_f64(uint64_t u)
{
    double x;
    __builtin_memcpy(&x, &u, sizeof(x));
    return x;
}

// ---- constants (exact binary values of the stated quantities) ----
constexpr double kLn2Hi = 0x1.62e42fefa39efp-1; // ln 2 = hi + lo
constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;
constexpr double kPio2a = 0x1.921fb54442d18p+0; // pi/2 = a + b + c, a and b
constexpr double kPio2b = 0x1.1a62633145c00p-54; // with 50 significant bits
constexpr double kPio2c = 0x1.b839a252049c1p-104;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;

// log1p(r) for a double-double |r| <= 2^-10.58, double-double result:
// r (1 - r/2 + r^2/3 - ... - r^9/10), the terms r^6/6 and beyond in double
// (the first omitted term is below 2^-116 relative)
__host__ __device__ inline ddv log1p_small(ddv r)
{
    double p = -0x1.999999999999ap-4;          // -1/10
    p = __builtin_fma(p, r.hi, 0x1.c71c71c71c71cp-4);  // 1/9
    p = __builtin_fma(p, r.hi, -0.125);                // -1/8
    p = __builtin_fma(p, r.hi, 0x1.2492492492492p-3);  // 1/7
    p = __builtin_fma(p, r.hi, -0x1.5555555555555p-3); // -1/6
    ddv s = add(ddv{0x1.999999999999ap-3, -0x1.999999999999ap-57},
        mul_d(r, p));                                   // 1/5
    s = add(ddv{-0.25, 0.0}, mul(r, s));
    s = add(ddv{0x1.5555555555555p-2, 0x1.5555555555555p-56}, mul(r, s)); // 1/3
    s = add(ddv{-0.5, 0.0}, mul(r, s));
    s = add(ddv{1.0, 0.0}, mul(r, s));
    return mul(r, s);
}

// log x for a positive normal double x, double-double result
__host__ __device__ inline ddv log_dd(double x, const double *log_c)
{
    const uint64_t u = as_u64(x);
    int e = (int)((u >> 52) & 0x7FF) - 1023;
    uint64_t m = (u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    if (u & 0x0008000000000000ull) { // f0 >= 1.5: halve it
        m -= 0x0010000000000000ull;
        e += 1;
    }
    const double f = as_f64(m); // [0.75, 1.5)
    const int j = (int)__builtin_rint(f * 1024.0);
    const double c = (double)j * 0x1p-10;
    const double d = f - c; // exact (Sterbenz)
    // r = d / c in double-double: the remainder d - q c is exact
    const double q = d / c;
    const ddv r = ddv{q, __builtin_fma(-q, c, d) / c};
    const ddv l1p = log1p_small(r);
    const ddv t{log_c[2 * (j - kLogJ0)], log_c[2 * (j - kLogJ0) + 1]};
    const double ed = (double)e;
    ddv el = two_prod(ed, kLn2Hi);
    el.lo = __builtin_fma(ed, kLn2Lo, el.lo);
    el = fast_two_sum(el.hi, el.lo);
    // e ln2 and log c_j share a sign or one is zero; log1p(r) is at most
    // half of |log c_j| (j != 1024) or the whole result (j == 1024, e == 0)
    return add_acc(add_acc(el, t), l1p);
}

// correctly rounded log x (x positive, normal)
__host__ __device__ inline double log_cr(double x, const double *log_c)
{
    const ddv v = log_dd(x, log_c);
    return v.hi + v.lo;
}

// log1p x for a finite double x > -1, double-double result.  |x| < 2^-11:
// the series.  Otherwise 1 + x = s + e exactly (two_sum; s >= 2^-53 is
// normal), log1p x = log s + log1p(e/s) with |e/s| <= 2^-53, so
// log1p(e/s) = t - t^2/2 to 2^-159, t = e/s in double-double (the
// remainder e - q s is exact).
__host__ __device__ inline ddv log1p_dd(double x, const double *log_c)
{
    if (__builtin_fabs(x) < 0x1p-11)
        return log1p_small(ddv{x, 0.0});
    const ddv s = two_sum(1.0, x);
    const ddv l = log_dd(s.hi, log_c);
    if (s.lo == 0.0)
        return l;
    const double q = s.lo / s.hi;
    const double tl = __builtin_fma(-q, s.hi, s.lo) / s.hi;
    return add(l, ddv{q, tl - 0.5 * (q * q)});
}

// correctly rounded log1p x (x > -1, finite)
__host__ __device__ inline double log1p_cr(double x, const double *log_c)
{
    if (x == 0.0)
        return x; // log1p(+-0) = +-0
    const ddv v = log1p_dd(x, log_c);
    return v.hi + v.lo;
}

// sin and cos of s, |s| <= 1/128 (double-double s)
__host__ __device__ inline void sincos_small(ddv s, ddv &sn, ddv &cs)
{
    const ddv z = mul(s, s);
    // sin s = s (1 - z/6 + z^2/120 - z^3/5040 + z^4/9! - z^5/11!)
    double qd = __builtin_fma(-0x1.ae64567f544e4p-26, z.hi, 0x1.71de3a556c734p-19);
    qd = __builtin_fma(qd, z.hi, -0x1.a01a01a01a01ap-13);
    ddv q = add(ddv{0x1.1111111111111p-7, 0x1.1111111111111p-63},
        mul_d(z, qd));
    q = add(ddv{-0x1.5555555555555p-3, -0x1.5555555555555p-57}, mul(z, q));
    q = add(ddv{1.0, 0.0}, mul(z, q));
    sn = mul(s, q);
    // cos s = 1 - z/2 + z^2/24 - z^3/720 + z^4/8! - z^5/10! + z^6/12!
    double cd = __builtin_fma(0x1.1eed8eff8d898p-29, z.hi, -0x1.27e4fb7789f5cp-22);
    cd = __builtin_fma(cd, z.hi, 0x1.a01a01a01a01ap-16);
    cd = __builtin_fma(cd, z.hi, -0x1.6c16c16c16c17p-10);
    ddv c = add(ddv{0x1.5555555555555p-5, 0x1.5555555555555p-59},
        mul_d(z, cd));
    c = add(ddv{-0.5, 0.0}, mul(z, c));
    cs = add(ddv{1.0, 0.0}, mul(z, c));
}

// sin x and cos x for 0 <= x <= 8, double-double results
__host__ __device__ inline void sincos_dd(double x, const double *sin_a,
    const double *cos_a, ddv &sn, ddv &cs)
{
    const double kd = __builtin_rint(x * kTwoOverPi);
    const int k = (int)kd;
    // r = x - k pi/2: k a is exact and x - k a is exact (Sterbenz), k b is
    // exact, k c carries the rest
    const double r1 = x - kd * kPio2a;
    ddv r = two_sum(r1, -kd * kPio2b);
    r = fast_two_sum(r.hi, __builtin_fma(-kd, kPio2c, r.lo));
    const bool negr = r.hi < 0.0;
    if (negr)
        r = neg(r);
    const int j = (int)__builtin_rint(r.hi * 64.0);
    ddv s_r, c_r;
    if (j == 0) {
        sincos_small(r, s_r, c_r);
    } else {
        const double a = (double)j * 0x1p-6;
        const ddv s = two_sum(r.hi - a, r.lo); // r.hi - a exact (Sterbenz)
        ddv ss, sc;
        sincos_small(s, ss, sc);
        const ddv sa{sin_a[2 * j], sin_a[2 * j + 1]};
        const ddv ca{cos_a[2 * j], cos_a[2 * j + 1]};
        s_r = add(mul(sa, sc), mul(ca, ss));
        c_r = add(mul(ca, sc), neg(mul(sa, ss)));
    }
    if (negr)
        s_r = neg(s_r);
    switch (k & 3) {
    case 0: sn = s_r; cs = c_r; break;
    case 1: sn = c_r; cs = neg(s_r); break;
    case 2: sn = neg(s_r); cs = neg(c_r); break;
    default: sn = neg(c_r); cs = s_r; break;
    }
}

// correctly rounded sin x and cos x, 0 <= x <= 8
__host__ __device__ inline void sincos_cr(double x, const double *sin_a,
    const double *cos_a, double &s, double &c)
{
    ddv sn, cs;
    sincos_dd(x, sin_a, cos_a, sn, cs);
    s = sn.hi + sn.lo;
    c = cs.hi + cs.lo;
}

// ---- tables, built on the host from series in double-double ----
// log c = 2 atanh(z), z = (c - 1)/(c + 1), |z| <= 1/5
inline ddv host_log(double c)
{
    if (c == 1.0)
        return ddv{0.0, 0.0};
    // z = (c - 1)/(c + 1): c - 1 and c + 1 are exact for c = j/1024
    const double num = c - 1.0, den = c + 1.0;
    const ddv z = div_d(ddv{num, 0.0}, den);
    const ddv z2 = mul(z, z);
    ddv term = z, sum = z;
    for (int k = 1; k < 40; ++k) {
        term = mul(term, z2);
        sum = add_acc(sum, div_d(term, (double)(2 * k + 1)));
    }
    return ddv{2.0 * sum.hi, 2.0 * sum.lo};
}

// sin a and cos a by their Taylor series, 0 <= a <= 0.8
inline void host_sincos(double a, ddv &s, ddv &c)
{
    const ddv x{a, 0.0};
    const ddv x2 = mul(x, x);
    ddv ts = x, tc{1.0, 0.0};
    s = ts;
    c = tc;
    for (int k = 1; k < 24; ++k) {
        ts = div_d(mul(ts, x2), (double)((2 * k) * (2 * k + 1)));
        tc = div_d(mul(tc, x2), (double)((2 * k - 1) * (2 * k)));
        s = add_acc(s, (k & 1) ? neg(ts) : ts);
        c = add_acc(c, (k & 1) ? neg(tc) : tc);
    }
}

inline void build_tables(Tables *t)
{
    for (int j = kLogJ0; j <= kLogJ1; ++j) {
        const ddv v = host_log((double)j / 1024.0);
        t->log_c[2 * (j - kLogJ0)] = v.hi;
        t->log_c[2 * (j - kLogJ0) + 1] = v.lo;
    }
    for (int j = 0; j < kTrigN; ++j) {
        ddv s, c;
        host_sincos((double)j / 64.0, s, c);
        t->sin_a[2 * j] = s.hi;
        t->sin_a[2 * j + 1] = s.lo;
        t->cos_a[2 * j] = c.hi;
        t->cos_a[2 * j + 1] = c.lo;
    }
}

} // namespace crm
} // namespace mckl_hip
