// primitives.hip -- device forms of the reference's Weight and resampling
// building blocks on caller-owned device arrays:
//   Weight::set_equal / set_log / add_log      core/weight.hpp:76-148,198-227
//   u01_rand_{sorted,stratified,systematic}    algorithm/resample.hpp:216-268
//   resample_trans_u01_rep                     algorithm/resample.hpp:374-407
//   resample_trans_rep_index                   algorithm/resample.hpp:416-466
//   ResampleAlgorithm + rep->index             algorithm/resample.hpp:494-576
//   StateMatrix::select (out of place)         core/state_matrix.hpp:193-265
// Prefix sums are double-double (rounded once to FP64), so they are the
// correctly rounded prefix sums except in astronomically rare ties; the
// reference's sequential FP64 sums carry up to ~sqrt(N) ulp of error, so a
// replication count can differ where a uniform lies that close to a
// cumulative-weight boundary.  Temporary buffers come from the stream-ordered
// allocator of the caller's stream.
#include "common.hpp"
#include "smc_math.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mckl_hip {

int rand_canonical_f64(mckl_hip_engine *e, size_t n, double *d_r, hipStream_t st);

namespace {

constexpr int kB = kGroup; // 256 threads, 8 items each: tiles of 2048

__global__ __launch_bounds__(kB) void k_weight_pre(int mode, uint64_t n,
    const double *v, double *w, double *pmax)
{
    // mode 0: w = v (set_log); mode 1: w = log(w) + v (add_log)
    __shared__ double sm[kB / 64];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    double m = -__builtin_huge_val();
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < n) {
            const double l = mode == 0 ? v[i] : __dadd_rn(v[i], log(w[i]));
            w[i] = l;
            m = m < l ? l : m;
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const double o = __shfl_xor(m, d, 64);
        m = m < o ? o : m;
    }
    if ((threadIdx.x & 63) == 0)
        sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sm[0];
        for (int i = 1; i < kB / 64; ++i)
            r = r < sm[i] ? sm[i] : r;
        pmax[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(kB) void k_weight_max(
    const double *pmax, uint32_t nb, double *lmax)
{
    __shared__ double sm[kB / 64];
    double m = -__builtin_huge_val();
    for (uint32_t i = threadIdx.x; i < nb; i += kB)
        m = m < pmax[i] ? pmax[i] : m;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const double o = __shfl_xor(m, d, 64);
        m = m < o ? o : m;
    }
    if ((threadIdx.x & 63) == 0)
        sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sm[0];
        for (int i = 1; i < kB / 64; ++i)
            r = r < sm[i] ? sm[i] : r;
        *lmax = r;
    }
}

// w = exp(w - lmax); block partial sums of w and w^2
__global__ __launch_bounds__(kB) void k_weight_exp(uint64_t n, double *w,
    const double *lmax, double *ps, double *ps2)
{
    __shared__ double sd[8];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    const double m = *lmax;
    double a = 0.0, b = 0.0;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < n) {
            const double e = exp(__dsub_rn(w[i], m));
            w[i] = e;
            a = __dadd_rn(a, e);
            b = __dadd_rn(b, __dmul_rn(e, e));
        }
    }
    group_sum2(a, b, threadIdx.x, sd);
    if (threadIdx.x == 0) {
        ps[blockIdx.x] = a;
        ps2[blockIdx.x] = b;
    }
}

// stats = {ESS, lmax + log(accw)}, scale = 1/accw
__global__ __launch_bounds__(kB) void k_weight_total(const double *ps,
    const double *ps2, uint32_t nb, const double *lmax, double *stats,
    double *scale)
{
    __shared__ double sd[8];
    double a, b;
    tiles_sum2(ps, ps2, (int)nb, threadIdx.x, sd, a, b);
    if (threadIdx.x == 0) {
        stats[0] = __ddiv_rn(__dmul_rn(a, a), b);
        stats[1] = __dadd_rn(*lmax, log(a));
        *scale = __ddiv_rn(1.0, a);
    }
}

__global__ void k_weight_scale(uint64_t n, double *w, const double *scale)
{
    const double s = *scale;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        w[i] = __dmul_rn(s, w[i]);
}

__global__ void k_weight_equal(uint64_t n, double *w, double *stats)
{
    const double v = __ddiv_rn(1.0, (double)n);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        w[i] = v;
    if (blockIdx.x == 0 && threadIdx.x == 0 && stats) {
        stats[0] = (double)n;
        stats[1] = 0.0;
    }
}

// ---- double-double scan of an array (tiles of 2048) ----
// term(i): w[i] (mode 0) or log(c[i]) / (N - i) (mode 1, sorted uniforms)
__device__ __forceinline__ double scan_term(int mode, const double *a,
    uint64_t i, uint64_t N)
{
    return mode == 0 ? a[i] : __ddiv_rn(log(a[i]), (double)(N - i));
}

__global__ __launch_bounds__(kB) void k_dd_tiles(int mode, uint64_t n,
    uint64_t N, const double *a, double *th, double *tl)
{
    __shared__ double sd[8];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    dd loc{0.0, 0.0};
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < n)
            loc = dd_add(loc, dd{scan_term(mode, a, i, N), 0.0});
    }
    dd tot;
    (void)group_excl_scan_dd(loc, threadIdx.x, sd, tot);
    if (threadIdx.x == 0) {
        th[blockIdx.x] = tot.hi;
        tl[blockIdx.x] = tot.lo;
    }
}

__global__ __launch_bounds__(kB) void k_dd_offsets(const double *th,
    const double *tl, uint32_t nt, double *oh, double *ol)
{
    __shared__ double sd[8];
    tiles_scan_dd(true, th, tl, (int)nt, oh, ol, threadIdx.x, sd);
}

// Inclusive prefix of element i, rounded: the last element of a tile takes
// the next tile's offset (the same convention as the SMC samplers).
__device__ __forceinline__ void tile_prefix(int mode, uint64_t n, uint64_t N,
    const double *a, const double *oh, const double *ol, double *P)
{
    __shared__ double sd[8];
    const uint32_t tile = blockIdx.x;
    const uint64_t i0 = (uint64_t)tile * kTile + threadIdx.x * kItems;
    const uint64_t tend = (uint64_t)(tile + 1) * kTile < n
        ? (uint64_t)(tile + 1) * kTile
        : n;
    double t[kItems];
    dd loc{0.0, 0.0};
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        t[k] = i < n ? scan_term(mode, a, i, N) : 0.0;
        loc = dd_add(loc, dd{t[k], 0.0});
    }
    dd tot;
    const dd off = group_excl_scan_dd(loc, threadIdx.x, sd, tot);
    const dd base = dd_add(dd{oh[tile], ol[tile]}, off);
    dd run{0.0, 0.0};
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        run = dd_add(run, dd{t[k], 0.0});
        P[k] = i == tend - 1 ? oh[tile + 1] : dd_add(base, run).hi;
    }
}

// sorted uniforms: r_j = 1 - exp(lm_j), lm = prefix of log(c_j)/(N-j)
// (resample.hpp:63-81)
__global__ __launch_bounds__(kB) void k_sorted_finish(uint64_t n,
    const double *c, const double *oh, const double *ol, double *r)
{
    double P[kItems];
    tile_prefix(1, n, n, c, oh, ol, P);
    __syncthreads();
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < n)
            r[i] = __dsub_rn(1.0, exp(P[k]));
    }
}

// #{j < M : u_j < S} for ascending u
__device__ __forceinline__ uint32_t count_below(
    const double *u, uint64_t M, double S)
{
    uint64_t lo = 0, hi = M;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (u[mid] < S)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (uint32_t)lo;
}

// cnt_i = #{u < S_i} for i < N-1, cnt_{N-1} = M
__global__ __launch_bounds__(kB) void k_counts(uint64_t N, uint64_t M,
    const double *w, const double *u, const double *oh, const double *ol,
    uint32_t *cnt)
{
    double P[kItems];
    tile_prefix(0, N, N, w, oh, ol, P);
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < N)
            cnt[i] = i == N - 1 ? (uint32_t)M : count_below(u, M, P[k]);
    }
}

__global__ void k_rep_small(uint64_t N, uint64_t M, uint32_t *rep)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        rep[i] = (N == 1 && i == 0) ? (uint32_t)M : 0u;
}

__global__ void k_rep_from_counts(uint64_t N, const uint32_t *cnt, uint32_t *rep)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        rep[i] = cnt[i] - (i == 0 ? 0u : cnt[i - 1]);
}

// ---- rep -> index, general M (resample.hpp:416-466) ----
// extras e_i = max(rep_i - 1, 0) for i < K, rep_i for i >= K (K = min(N,M));
// demands: zero slots d < K in order, then slots K..M-1.
__device__ __forceinline__ uint32_t extras(uint32_t rep, uint64_t i, uint64_t K)
{
    return i < K ? (rep > 1 ? rep - 1 : 0u) : rep;
}

__global__ __launch_bounds__(kB) void k_ri_tiles(uint64_t N, uint64_t K,
    const uint32_t *rep, uint32_t *te, uint32_t *tz)
{
    __shared__ uint32_t su[8];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    uint32_t es = 0, zs = 0;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < N) {
            es += extras(rep[i], i, K);
            zs += (i < K && rep[i] == 0) ? 1u : 0u;
        }
    }
    uint32_t a, b;
    (void)group_excl_scan_u(es, threadIdx.x, su, a);
    (void)group_excl_scan_u(zs, threadIdx.x, su + 4, b);
    if (threadIdx.x == 0) {
        te[blockIdx.x] = a;
        tz[blockIdx.x] = b;
    }
}

__global__ __launch_bounds__(kB) void k_ri_offsets(const uint32_t *te,
    const uint32_t *tz, uint32_t nt, uint32_t *teo, uint32_t *tzo)
{
    __shared__ uint32_t su[8];
    tiles_scan_u2(true, te, tz, (int)nt, teo, tzo, threadIdx.x, su);
}

__global__ __launch_bounds__(kB) void k_ri_write(uint64_t N, uint64_t K,
    const uint32_t *rep, const uint32_t *teo, const uint32_t *tzo,
    uint32_t *Eincl, uint32_t *idx)
{
    __shared__ uint32_t su[8];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    uint32_t es = 0, zs = 0;
    uint32_t r[kItems];
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        r[k] = i < N ? rep[i] : 0u;
        if (i < N) {
            es += extras(r[k], i, K);
            zs += (i < K && r[k] == 0) ? 1u : 0u;
        }
    }
    uint32_t a, b;
    uint32_t ex = group_excl_scan_u(es, threadIdx.x, su, a) + teo[blockIdx.x];
    uint32_t zx = group_excl_scan_u(zs, threadIdx.x, su + 4, b) + tzo[blockIdx.x];
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < N) {
            ex += extras(r[k], i, K);
            Eincl[i] = ex;
            if (i < K) {
                if (r[k] > 0) {
                    idx[i] = (uint32_t)i;
                } else {
                    idx[i] = 0x80000000u | zx;
                    ++zx;
                }
            }
        }
    }
}

__device__ __forceinline__ uint32_t ub(const uint32_t *E, uint64_t n, uint32_t z)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (E[mid] > z)
            hi = mid;
        else
            lo = mid + 1;
    }
    return (uint32_t)lo;
}

__global__ void k_ri_resolve(uint64_t N, uint64_t K, uint64_t M,
    const uint32_t *Eincl, const uint32_t *tzo, uint32_t ntiles, uint32_t *idx)
{
    const uint32_t ztot = tzo[ntiles];
    for (uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; d < M;
         d += (uint64_t)gridDim.x * blockDim.x) {
        if (d < K) {
            const uint32_t v = idx[d];
            if (v & 0x80000000u)
                idx[d] = ub(Eincl, N, v & 0x7FFFFFFFu);
        } else {
            idx[d] = ub(Eincl, N, ztot + (uint32_t)(d - K));
        }
    }
}

__global__ void k_systematic(uint64_t n, const double *u01, double *r)
{
    const double delta = __ddiv_rn(1.0, (double)n);
    const double u = __dmul_rn(*u01, delta);
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x)
        r[j] = __dadd_rn(__dmul_rn((double)j, delta), u);
}

__global__ void k_stratified(uint64_t n, double *r)
{
    const double delta = __ddiv_rn(1.0, (double)n);
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (uint64_t)gridDim.x * blockDim.x)
        r[j] = __dmul_rn(delta, __dadd_rn(r[j], (double)j));
}

template <typename T>
__global__ void k_select(uint64_t M, uint64_t dim, const uint32_t *idx,
    const T *src, T *dst)
{
    const uint64_t total = M * dim;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
         q < total; q += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j = q / dim, c = q % dim;
        dst[q] = src[(uint64_t)idx[j] * dim + c];
    }
}

uint32_t grid_for(uint64_t n, uint32_t cap = 8192)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256), cap));
}

// RAII for stream-ordered scratch
struct Scratch {
    hipStream_t st;
    void *p = nullptr;
    explicit Scratch(hipStream_t s) : st(s) {}
    int alloc(size_t bytes)
    {
        MCKL_HIP_TRY(hipMallocAsync(&p, std::max<size_t>(bytes, 256), st));
        return MCKL_HIP_OK;
    }
    ~Scratch()
    {
        if (p)
            (void)hipFreeAsync(p, st);
    }
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

int weight_normalize(int mode, size_t n, const double *v, double *w,
    double *stats, hipStream_t st)
{
    const uint32_t nt = (uint32_t)ceil_div(n, kTile);
    Scratch s(st);
    const size_t o1 = align256(nt * sizeof(double));
    int rc = s.alloc(3 * o1 + 256);
    if (rc)
        return rc;
    char *base = static_cast<char *>(s.p);
    double *pmax = reinterpret_cast<double *>(base);
    double *ps = reinterpret_cast<double *>(base + o1);
    double *ps2 = reinterpret_cast<double *>(base + 2 * o1);
    double *lmax = reinterpret_cast<double *>(base + 3 * o1);
    double *scale = lmax + 1;
    k_weight_pre<<<nt, kB, 0, st>>>(mode, n, v, w, pmax);
    k_weight_max<<<1, kB, 0, st>>>(pmax, nt, lmax);
    k_weight_exp<<<nt, kB, 0, st>>>(n, w, lmax, ps, ps2);
    k_weight_total<<<1, kB, 0, st>>>(ps, ps2, nt, lmax, stats, scale);
    k_weight_scale<<<grid_for(n), 256, 0, st>>>(n, w, scale);
    MCKL_HIP_CHECK_LAUNCH("weight kernels");
    return MCKL_HIP_OK;
}

// dd prefix offsets of the scan terms; returns offsets in oh/ol (nt+1)
int dd_offsets(int mode, size_t n, const double *a, double *th, double *tl,
    double *oh, double *ol, hipStream_t st)
{
    const uint32_t nt = (uint32_t)ceil_div(n, kTile);
    k_dd_tiles<<<nt, kB, 0, st>>>(mode, n, n, a, th, tl);
    k_dd_offsets<<<1, kB, 0, st>>>(th, tl, nt, oh, ol);
    MCKL_HIP_CHECK_LAUNCH("dd scan kernels");
    return MCKL_HIP_OK;
}

int sequence(int scheme, mckl_hip_engine *e, size_t n, double *r, hipStream_t st)
{
    if (scheme == MCKL_HIP_SYSTEMATIC) {
        // one U01Canonical<double> draw (two words), resample.hpp:258-268
        Scratch s(st);
        int rc = s.alloc(256);
        if (rc)
            return rc;
        double *u01 = static_cast<double *>(s.p);
        rc = rand_canonical_f64(e, 1, u01, st);
        if (rc)
            return rc;
        if (n == 0)
            return MCKL_HIP_OK;
        k_systematic<<<grid_for(n), 256, 0, st>>>(n, u01, r);
        MCKL_HIP_CHECK_LAUNCH("k_systematic");
        return MCKL_HIP_OK;
    }
    if (n == 0)
        return MCKL_HIP_OK;
    int rc = rand_canonical_f64(e, n, r, st);
    if (rc)
        return rc;
    if (scheme == MCKL_HIP_STRATIFIED) {
        k_stratified<<<grid_for(n), 256, 0, st>>>(n, r);
        MCKL_HIP_CHECK_LAUNCH("k_stratified");
        return MCKL_HIP_OK;
    }
    // sorted (multinomial)
    const uint32_t nt = (uint32_t)ceil_div(n, kTile);
    Scratch s(st);
    const size_t o = align256((nt + 1) * sizeof(double));
    rc = s.alloc(4 * o + align256(n * sizeof(double)));
    if (rc)
        return rc;
    char *base = static_cast<char *>(s.p);
    double *th = reinterpret_cast<double *>(base);
    double *tl = reinterpret_cast<double *>(base + o);
    double *oh = reinterpret_cast<double *>(base + 2 * o);
    double *ol = reinterpret_cast<double *>(base + 3 * o);
    double *c = reinterpret_cast<double *>(base + 4 * o);
    MCKL_HIP_TRY(hipMemcpyAsync(c, r, n * sizeof(double),
        hipMemcpyDeviceToDevice, st));
    rc = dd_offsets(1, n, c, th, tl, oh, ol, st);
    if (rc)
        return rc;
    k_sorted_finish<<<nt, kB, 0, st>>>(n, c, oh, ol, r);
    MCKL_HIP_CHECK_LAUNCH("k_sorted_finish");
    return MCKL_HIP_OK;
}

int trans_u01_rep(size_t N, size_t M, const double *w, const double *u,
    uint32_t *rep, hipStream_t st)
{
    if (N == 0)
        return MCKL_HIP_OK;
    if (N == 1 || M == 0) {
        // resample.hpp:387-397: rep = {M} for N == 1, zeros for M == 0
        k_rep_small<<<grid_for(N), 256, 0, st>>>(N, M, rep);
        MCKL_HIP_CHECK_LAUNCH("k_rep_small");
        return MCKL_HIP_OK;
    }
    const uint32_t nt = (uint32_t)ceil_div(N, kTile);
    Scratch s(st);
    const size_t o = align256((nt + 1) * sizeof(double));
    int rc = s.alloc(4 * o + align256(N * sizeof(uint32_t)));
    if (rc)
        return rc;
    char *base = static_cast<char *>(s.p);
    double *th = reinterpret_cast<double *>(base);
    double *tl = reinterpret_cast<double *>(base + o);
    double *oh = reinterpret_cast<double *>(base + 2 * o);
    double *ol = reinterpret_cast<double *>(base + 3 * o);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(base + 4 * o);
    rc = dd_offsets(0, N, w, th, tl, oh, ol, st);
    if (rc)
        return rc;
    k_counts<<<nt, kB, 0, st>>>(N, M, w, u, oh, ol, cnt);
    k_rep_from_counts<<<grid_for(N), 256, 0, st>>>(N, cnt, rep);
    MCKL_HIP_CHECK_LAUNCH("replication kernels");
    return MCKL_HIP_OK;
}

int rep_index(size_t N, size_t M, const uint32_t *rep, uint32_t *idx,
    hipStream_t st)
{
    if (N == 0 || M == 0)
        return MCKL_HIP_OK;
    const size_t K = std::min(N, M);
    const uint32_t nt = (uint32_t)ceil_div(N, kTile);
    Scratch s(st);
    const size_t o = align256((nt + 1) * sizeof(uint32_t));
    int rc = s.alloc(4 * o + align256(N * sizeof(uint32_t)));
    if (rc)
        return rc;
    char *base = static_cast<char *>(s.p);
    uint32_t *te = reinterpret_cast<uint32_t *>(base);
    uint32_t *tz = reinterpret_cast<uint32_t *>(base + o);
    uint32_t *teo = reinterpret_cast<uint32_t *>(base + 2 * o);
    uint32_t *tzo = reinterpret_cast<uint32_t *>(base + 3 * o);
    uint32_t *E = reinterpret_cast<uint32_t *>(base + 4 * o);
    k_ri_tiles<<<nt, kB, 0, st>>>(N, K, rep, te, tz);
    k_ri_offsets<<<1, kB, 0, st>>>(te, tz, nt, teo, tzo);
    k_ri_write<<<nt, kB, 0, st>>>(N, K, rep, teo, tzo, E, idx);
    k_ri_resolve<<<grid_for(M), 256, 0, st>>>(N, K, M, E, tzo, nt, idx);
    MCKL_HIP_CHECK_LAUNCH("rep->index kernels");
    return MCKL_HIP_OK;
}

int check_scheme(int scheme)
{
    if (scheme != MCKL_HIP_MULTINOMIAL && scheme != MCKL_HIP_STRATIFIED &&
        scheme != MCKL_HIP_SYSTEMATIC)
        return set_error(MCKL_HIP_EINVAL, "unknown resampling scheme %d", scheme);
    return MCKL_HIP_OK;
}

// ---- residual family: resample_trans_residual (resample.hpp:333-364) ----
// resid_i = modf(M W_i, &integ_i); tile sums: resid in double-double,
// integ exactly
__global__ __launch_bounds__(kB) void k_residual_tiles(uint64_t N, double Md,
    const double *w, double *resid, uint32_t *integ, double *th, double *tl,
    unsigned long long *ti)
{
    __shared__ __align__(16) double sd[16];
    __shared__ unsigned long long su[4];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    dd loc{0.0, 0.0};
    unsigned long long isum = 0;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < N) {
            double ip;
            const double r = modf(__dmul_rn(Md, w[i]), &ip);
            resid[i] = r;
            integ[i] = (uint32_t)ip;
            loc = dd_add(loc, dd{r, 0.0});
            isum += (unsigned long long)ip;
        }
    }
    dd tot;
    (void)group_excl_scan_dd(loc, threadIdx.x, sd, tot);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        isum += __shfl_xor(isum, d, 64);
    if ((threadIdx.x & 63) == 0)
        su[threadIdx.x >> 6] = isum;
    __syncthreads();
    if (threadIdx.x == 0) {
        th[blockIdx.x] = tot.hi;
        tl[blockIdx.x] = tot.lo;
        ti[blockIdx.x] = su[0] + su[1] + su[2] + su[3];
    }
}

// out[0] = 1 / sum(resid) (the sum rounded once from double-double),
// out[1] = M - sum(integ) (as a double-exact integer < 2^31)
__global__ __launch_bounds__(kB) void k_residual_total(const double *th,
    const double *tl, const unsigned long long *ti, uint32_t nt, uint64_t M,
    double *mul, unsigned long long *R)
{
    __shared__ __align__(16) double sd[16];
    __shared__ unsigned long long su[4];
    const int gt = threadIdx.x;
    const int chunk = ((int)nt + kB - 1) / kB;
    const int b = gt * chunk;
    const int e = b + chunk < (int)nt ? b + chunk : (int)nt;
    dd loc{0.0, 0.0};
    unsigned long long isum = 0;
    for (int i = b; i < e; ++i) {
        loc = dd_add(loc, dd{th[i], tl[i]});
        isum += ti[i];
    }
    dd tot;
    (void)group_excl_scan_dd(loc, gt, sd, tot);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        isum += __shfl_xor(isum, d, 64);
    if ((gt & 63) == 0)
        su[gt >> 6] = isum;
    __syncthreads();
    if (gt == 0) {
        *mul = __ddiv_rn(1.0, tot.hi);
        *R = M - (su[0] + su[1] + su[2] + su[3]);
    }
}

__global__ void k_residual_scale(uint64_t N, double *resid, const double *mul)
{
    const double s = *mul;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        resid[i] = __dmul_rn(resid[i], s);
}

__global__ void k_add_integ(uint64_t N, uint32_t *rep, const uint32_t *integ)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        rep[i] += integ[i];
}

// ResampleAlgorithm<U01Seq, true>::eval (resample.hpp:555-576): residuals,
// R uniforms of the base scheme over them, plus the integral parts.  R is
// read back to the host (the uniform sequence advances the engine by R).
int residual_rep(int base, size_t N, size_t M, mckl_hip_engine *e,
    const double *w, uint32_t *rep, hipStream_t st)
{
    const uint32_t nt = (uint32_t)ceil_div(N, kTile);
    Scratch s(st);
    const size_t o = align256(nt * sizeof(double));
    const size_t on = align256(N * sizeof(double));
    const size_t oi = align256(N * sizeof(uint32_t));
    int rc = s.alloc(3 * o + on + oi + 512);
    if (rc)
        return rc;
    char *base_p = static_cast<char *>(s.p);
    double *th = reinterpret_cast<double *>(base_p);
    double *tl = reinterpret_cast<double *>(base_p + o);
    unsigned long long *ti =
        reinterpret_cast<unsigned long long *>(base_p + 2 * o);
    double *resid = reinterpret_cast<double *>(base_p + 3 * o);
    uint32_t *integ = reinterpret_cast<uint32_t *>(base_p + 3 * o + on);
    double *mul = reinterpret_cast<double *>(base_p + 3 * o + on + oi);
    unsigned long long *dR = reinterpret_cast<unsigned long long *>(
        base_p + 3 * o + on + oi + 256);
    k_residual_tiles<<<nt, kB, 0, st>>>(
        N, (double)M, w, resid, integ, th, tl, ti);
    k_residual_total<<<1, kB, 0, st>>>(th, tl, ti, nt, M, mul, dR);
    k_residual_scale<<<grid_for(N), 256, 0, st>>>(N, resid, mul);
    MCKL_HIP_CHECK_LAUNCH("residual kernels");
    unsigned long long R = 0;
    MCKL_HIP_TRY(hipMemcpyAsync(&R, dR, sizeof(R), hipMemcpyDeviceToHost, st));
    MCKL_HIP_TRY(hipStreamSynchronize(st));
    if (R > M)
        return set_error(MCKL_HIP_EINVAL,
            "weights do not sum to one (integral parts exceed M)");
    Scratch su(st);
    rc = su.alloc(align256(std::max<size_t>(R, 1) * sizeof(double)));
    if (rc)
        return rc;
    double *u = static_cast<double *>(su.p);
    rc = sequence(base, e, R, u, st);
    if (rc)
        return rc;
    rc = trans_u01_rep(N, R, resid, u, rep, st);
    if (rc)
        return rc;
    k_add_integ<<<grid_for(N), 256, 0, st>>>(N, rep, integ);
    MCKL_HIP_CHECK_LAUNCH("k_add_integ");
    return MCKL_HIP_OK;
}

// ---- SMCEstimator::estimate weighted sums (algorithm/smc.hpp:89-118) ----
// r_j = sum_i w_i u(i, j).  Each product is split exactly (two_prod) and
// summed in double-double, per tile of 2048 rows and then over tiles in a
// fixed order, so r_j is the exactly rounded sum up to ~2^-104 relative to
// sum |w_i u(i, j)|, whatever the grid.
constexpr int kEstCols = 8; // columns per workgroup

template <typename U>
__global__ __launch_bounds__(kB) void k_est_tiles(uint64_t n, uint32_t dim,
    int row_major, const U *__restrict__ u, const double *__restrict__ w,
    double *phi, double *plo)
{
    __shared__ double sm[2 * kEstCols * (kB / 64)];
    const uint32_t j0 = blockIdx.y * kEstCols;
    const uint32_t nj = min((uint32_t)kEstCols, dim - j0);
    dd acc[kEstCols];
#pragma unroll
    for (int j = 0; j < kEstCols; ++j)
        acc[j] = dd{0.0, 0.0};
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x;
#pragma unroll 2
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + (uint64_t)k * kB;
        if (i >= n)
            break;
        const double wi = w[i];
#pragma unroll
        for (int j = 0; j < kEstCols; ++j) {
            if ((uint32_t)j < nj) {
                const double v = row_major ? (double)u[i * dim + j0 + j]
                                           : (double)u[(uint64_t)(j0 + j) * n + i];
                acc[j] = dd_add(acc[j], two_prod(wi, v));
            }
        }
    }
    const int wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kEstCols; ++j) {
        const dd t = wave_sum_dd(acc[j]);
        if ((threadIdx.x & 63) == 0) {
            sm[2 * (j * (kB / 64) + wv)] = t.hi;
            sm[2 * (j * (kB / 64) + wv) + 1] = t.lo;
        }
    }
    __syncthreads();
    if (threadIdx.x < nj) {
        const int j = threadIdx.x;
        dd t{0.0, 0.0};
        for (int q = 0; q < kB / 64; ++q)
            t = dd_add(t, dd{sm[2 * (j * (kB / 64) + q)],
                               sm[2 * (j * (kB / 64) + q) + 1]});
        const uint64_t o = (uint64_t)(j0 + j) * gridDim.x + blockIdx.x;
        phi[o] = t.hi;
        plo[o] = t.lo;
    }
}

// one workgroup per column: dd sum of the column's tile partials in tile
// order (chunked per thread, then across the group)
__global__ __launch_bounds__(kB) void k_est_final(uint32_t nt,
    const double *phi, const double *plo, double *r)
{
    __shared__ double sm[2 * (kB / 64)];
    const uint64_t base = (uint64_t)blockIdx.x * nt;
    dd t{0.0, 0.0};
    for (uint32_t q = threadIdx.x; q < nt; q += kB)
        t = dd_add(t, dd{phi[base + q], plo[base + q]});
    t = wave_sum_dd(t);
    if ((threadIdx.x & 63) == 0) {
        sm[2 * (threadIdx.x >> 6)] = t.hi;
        sm[2 * (threadIdx.x >> 6) + 1] = t.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        dd s{0.0, 0.0};
        for (int q = 0; q < kB / 64; ++q)
            s = dd_add(s, dd{sm[2 * q], sm[2 * q + 1]});
        r[blockIdx.x] = __dadd_rn(s.hi, s.lo);
    }
}

template <typename T>
int select_impl(size_t M, size_t dim, const uint32_t *idx, const T *src,
    T *dst, hipStream_t st)
{
    if (M == 0 || dim == 0)
        return MCKL_HIP_OK;
    if (!idx || !src || !dst)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (src == dst)
        return set_error(MCKL_HIP_EINVAL, "select is out of place");
    if (dim > ((uint64_t)1 << 40) / M)
        return set_error(MCKL_HIP_EINVAL, "select: size out of range");
    k_select<T><<<grid_for(M * dim), 256, 0, st>>>(M, dim, idx, src, dst);
    MCKL_HIP_CHECK_LAUNCH("k_select");
    return MCKL_HIP_OK;
}

template <typename U>
int estimate_impl(size_t n, size_t dim, int layout, const U *u,
    const double *w, double *r, hipStream_t st)
{
    if (layout != MCKL_HIP_ROW_MAJOR && layout != MCKL_HIP_COL_MAJOR)
        return set_error(MCKL_HIP_EINVAL, "unknown matrix layout %d", layout);
    if (dim == 0)
        return MCKL_HIP_OK;
    if (!r || (n && (!u || !w)))
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (dim > 65535u * kEstCols || n > ((uint64_t)1 << 40))
        return set_error(MCKL_HIP_EINVAL, "estimate: size out of range");
    const uint32_t nt = (uint32_t)std::max<uint64_t>(1, ceil_div(n, kTile));
    Scratch s(st);
    const size_t half = align256(nt * dim * sizeof(double));
    int rc = s.alloc(2 * half);
    if (rc)
        return rc;
    double *phi = static_cast<double *>(s.p);
    double *plo = reinterpret_cast<double *>(static_cast<char *>(s.p) + half);
    const dim3 grid(nt, (uint32_t)ceil_div(dim, kEstCols));
    k_est_tiles<U><<<grid, kB, 0, st>>>(
        n, (uint32_t)dim, layout == MCKL_HIP_ROW_MAJOR, u, w, phi, plo);
    MCKL_HIP_CHECK_LAUNCH("k_est_tiles");
    k_est_final<<<(uint32_t)dim, kB, 0, st>>>(nt, phi, plo, r);
    MCKL_HIP_CHECK_LAUNCH("k_est_final");
    return MCKL_HIP_OK;
}

} // namespace
} // namespace mckl_hip

using namespace mckl_hip;

extern "C" {

int mckl_hip_weight_set_equal(size_t n, double *d_w, double *d_stats,
    void *stream)
{
    if (n == 0)
        return MCKL_HIP_OK;
    if (!d_w)
        return set_error(MCKL_HIP_EINVAL, "NULL weights");
    k_weight_equal<<<grid_for(n), 256, 0, as_stream(stream)>>>(n, d_w, d_stats);
    MCKL_HIP_CHECK_LAUNCH("k_weight_equal");
    return MCKL_HIP_OK;
}

int mckl_hip_weight_set_log(size_t n, const double *d_v, double *d_w,
    double *d_stats, void *stream)
{
    if (n == 0)
        return MCKL_HIP_OK;
    if (!d_v || !d_w || !d_stats)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    return weight_normalize(0, n, d_v, d_w, d_stats, as_stream(stream));
}

int mckl_hip_weight_add_log(size_t n, const double *d_v, double *d_w,
    double *d_stats, void *stream)
{
    if (n == 0)
        return MCKL_HIP_OK;
    if (!d_v || !d_w || !d_stats)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    return weight_normalize(1, n, d_v, d_w, d_stats, as_stream(stream));
}

int mckl_hip_u01_rand_sequence(
    int scheme, mckl_hip_engine *e, size_t n, double *d_r, void *stream)
{
    int rc = check_scheme(scheme);
    if (rc)
        return rc;
    if (!e || (n && !d_r))
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (e->rbits != 32)
        return set_error(MCKL_HIP_EINVAL, "uniform sequences use 32-bit engines");
    return sequence(scheme, e, n, d_r, as_stream(stream));
}

int mckl_hip_resample_trans_u01_rep(size_t N, size_t M, const double *d_w,
    const double *d_u01, uint32_t *d_rep, void *stream)
{
    if (N >= 0x7FFFFFFFu || M >= 0x7FFFFFFFu)
        return set_error(MCKL_HIP_EINVAL, "sizes must be < 2^31");
    if (N && (!d_w || !d_rep || (M && !d_u01)))
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    return trans_u01_rep(N, M, d_w, d_u01, d_rep, as_stream(stream));
}

int mckl_hip_resample_trans_rep_index(size_t N, size_t M,
    const uint32_t *d_rep, uint32_t *d_idx, void *stream)
{
    if (N >= 0x7FFFFFFFu || M >= 0x7FFFFFFFu)
        return set_error(MCKL_HIP_EINVAL, "sizes must be < 2^31");
    if (N && M && (!d_rep || !d_idx))
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    return rep_index(N, M, d_rep, d_idx, as_stream(stream));
}

int mckl_hip_resample(int scheme, size_t N, size_t M, mckl_hip_engine *e,
    const double *d_w, uint32_t *d_rep, uint32_t *d_idx, void *stream)
{
    const bool residual = scheme >= MCKL_HIP_RESIDUAL &&
        scheme <= MCKL_HIP_RESIDUAL_SYSTEMATIC;
    int rc = check_scheme(residual ? scheme - MCKL_HIP_RESIDUAL : scheme);
    if (rc)
        return rc;
    if (!e || !d_rep)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (N >= 0x7FFFFFFFu || M >= 0x7FFFFFFFu)
        return set_error(MCKL_HIP_EINVAL, "sizes must be < 2^31");
    hipStream_t st = as_stream(stream);
    if (residual) {
        if (N == 0)
            return MCKL_HIP_OK;
        rc = residual_rep(scheme - MCKL_HIP_RESIDUAL, N, M, e, d_w, d_rep, st);
        if (rc)
            return rc;
        if (d_idx)
            rc = rep_index(N, M, d_rep, d_idx, st);
        return rc;
    }
    Scratch s(st);
    rc = s.alloc(align256(std::max<size_t>(M, 1) * sizeof(double)));
    if (rc)
        return rc;
    double *u = static_cast<double *>(s.p);
    rc = sequence(scheme, e, M, u, st);
    if (rc)
        return rc;
    rc = trans_u01_rep(N, M, d_w, u, d_rep, st);
    if (rc)
        return rc;
    if (d_idx)
        rc = rep_index(N, M, d_rep, d_idx, st);
    return rc;
}

int mckl_hip_estimate_f64(size_t n, size_t dim, int layout, const double *d_u,
    const double *d_w, double *d_r, void *stream)
{
    return estimate_impl(n, dim, layout, d_u, d_w, d_r, as_stream(stream));
}

int mckl_hip_estimate_f32(size_t n, size_t dim, int layout, const float *d_u,
    const double *d_w, double *d_r, void *stream)
{
    return estimate_impl(n, dim, layout, d_u, d_w, d_r, as_stream(stream));
}

int mckl_hip_select_f32(size_t M, size_t dim, const uint32_t *d_idx,
    const float *d_src, float *d_dst, void *stream)
{
    return select_impl(M, dim, d_idx, d_src, d_dst, as_stream(stream));
}

int mckl_hip_select_f64(size_t M, size_t dim, const uint32_t *d_idx,
    const double *d_src, double *d_dst, void *stream)
{
    return select_impl(M, dim, d_idx, d_src, d_dst, as_stream(stream));
}

} // extern "C"
