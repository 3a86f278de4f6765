// Host build of mckl_amd/csrc/crmath.hpp (the device's correctly rounded
// log/log1p/sin/cos) for the CPU tests: the same source, compiled by g++ with
// -ffp-contract=off, so its results can be compared with quad precision
// without a GPU.  Test infrastructure only.
#define __host__
#define __device__
#include "../../mckl_amd/csrc/crmath.hpp"

#include <cstddef>

using namespace mckl_hip::crm;

static Tables g_tables;
static bool g_built = false;

static const Tables &tables()
{
    if (!g_built) {
        build_tables(&g_tables);
        g_built = true;
    }
    return g_tables;
}

extern "C" {

void crm_host_log(size_t n, const double *x, double *y)
{
    const Tables &t = tables();
    for (size_t i = 0; i < n; ++i)
        y[i] = log_cr(x[i], t.log_c);
}

void crm_host_log1p(size_t n, const double *x, double *y)
{
    const Tables &t = tables();
    for (size_t i = 0; i < n; ++i)
        y[i] = log1p_cr(x[i], t.log_c);
}

void crm_host_sincos(size_t n, const double *x, double *s, double *c)
{
    const Tables &t = tables();
    for (size_t i = 0; i < n; ++i)
        sincos_cr(x[i], t.sin_a, t.cos_a, s[i], c[i]);
}

// the double-double results before rounding (hi, lo interleaved)
void crm_host_log_dd(size_t n, const double *x, double *y)
{
    const Tables &t = tables();
    for (size_t i = 0; i < n; ++i) {
        const ddv v = log_dd(x[i], t.log_c);
        y[2 * i] = v.hi;
        y[2 * i + 1] = v.lo;
    }
}

void crm_host_sincos_dd(size_t n, const double *x, double *s, double *c)
{
    const Tables &t = tables();
    for (size_t i = 0; i < n; ++i) {
        ddv a, b;
        sincos_dd(x[i], t.sin_a, t.cos_a, a, b);
        s[2 * i] = a.hi;
        s[2 * i + 1] = a.lo;
        c[2 * i] = b.hi;
        c[2 * i + 1] = b.lo;
    }
}

void crm_host_tables(double *log_c, double *sin_a, double *cos_a)
{
    const Tables &t = tables();
    for (int i = 0; i < 2 * kLogN; ++i)
        log_c[i] = t.log_c[i];
    for (int i = 0; i < 2 * kTrigN; ++i) {
        sin_a[i] = t.sin_a[i];
        cos_a[i] = t.cos_a[i];
    }
}

int crm_host_sizes(int *nlog, int *ntrig, int *j0)
{
    *nlog = kLogN;
    *ntrig = kTrigN;
    *j0 = kLogJ0;
    return 0;
}

} // extern "C"
