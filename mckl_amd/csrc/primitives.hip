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
#include "smc_index.hpp"
#include "smc_math.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace mckl_hip {

int rand_canonical_f64(mckl_hip_engine *e, size_t n, double *d_r, hipStream_t st);

namespace {

constexpr int kB = kGroup; // 256 threads, 8 items each: tiles of 2048

// A thread's eight elements i0 .. i0+7: 16-byte accesses when all eight
// are below n and p + i0 is 16-byte aligned, else element by element
// (fill past n)
__device__ __forceinline__ bool vec_ok(const void *p, uint64_t i0, uint64_t n)
{
    return i0 + kItems <= n && (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

__device__ __forceinline__ void ld8(const double *p, uint64_t i0, uint64_t n,
    double (&v)[kItems], double fill)
{
    if (vec_ok(p + i0, i0, n)) {
#pragma unroll
        for (int k = 0; k < kItems; k += 2) {
            const double2 a = *reinterpret_cast<const double2 *>(p + i0 + k);
            v[k] = a.x;
            v[k + 1] = a.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            v[k] = i0 + k < n ? p[i0 + k] : fill;
    }
}

__device__ __forceinline__ void st8(double *p, uint64_t i0, uint64_t n,
    const double (&v)[kItems])
{
    if (vec_ok(p + i0, i0, n)) {
#pragma unroll
        for (int k = 0; k < kItems; k += 2)
            *reinterpret_cast<double2 *>(p + i0 + k) = double2{v[k], v[k + 1]};
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (i0 + k < n)
                p[i0 + k] = v[k];
    }
}

__device__ __forceinline__ void ld8(const uint32_t *p, uint64_t i0,
    uint64_t n, uint32_t (&v)[kItems], uint32_t fill)
{
    if (vec_ok(p + i0, i0, n)) {
#pragma unroll
        for (int k = 0; k < kItems; k += 4) {
            const uint4 a = *reinterpret_cast<const uint4 *>(p + i0 + k);
            v[k] = a.x;
            v[k + 1] = a.y;
            v[k + 2] = a.z;
            v[k + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            v[k] = i0 + k < n ? p[i0 + k] : fill;
    }
}

__device__ __forceinline__ void st8(uint32_t *p, uint64_t i0, uint64_t n,
    const uint32_t (&v)[kItems])
{
    if (vec_ok(p + i0, i0, n)) {
#pragma unroll
        for (int k = 0; k < kItems; k += 4)
            *reinterpret_cast<uint4 *>(p + i0 + k) =
                make_uint4(v[k], v[k + 1], v[k + 2], v[k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (i0 + k < n)
                p[i0 + k] = v[k];
    }
}

__global__ __launch_bounds__(kB) void k_weight_pre(int mode, uint64_t n,
    const double *v, double *w, double *pmax)
{
    // mode 0: w = v (set_log); mode 1: w = log(w) + v (add_log)
    __shared__ double sm[kB / 64];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    double m = -__builtin_huge_val();
    double a[kItems], b[kItems];
    ld8(v, i0, n, a, 0.0);
    if (mode != 0)
        ld8(w, i0, n, b, 1.0);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < n) {
            const double l = mode == 0 ? a[k] : __dadd_rn(a[k], log(b[k]));
            a[k] = l;
            m = m < l ? l : m;
        }
    }
    st8(w, i0, n, a);
    m = wave_max_d(m);
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

// max over the n block maxima (every thread gets it)
__device__ __forceinline__ double blocks_max(const double *pmax, uint32_t n,
    double *sm)
{
    double m = -__builtin_huge_val();
#pragma unroll 4
    for (uint32_t i = threadIdx.x; i < n; i += kB)
        m = m < pmax[i] ? pmax[i] : m;
    m = wave_max_d(m);
    if ((threadIdx.x & 63) == 0)
        sm[threadIdx.x >> 6] = m;
    __syncthreads();
    double r = sm[0];
    for (int i = 1; i < kB / 64; ++i)
        r = r < sm[i] ? sm[i] : r;
    __syncthreads();
    return r;
}

// Many tiles: one workgroup reduces the block maxima (gmax[0]), or later
// the partial sums (gsum[0..1], and stats), and the other kernels read the
// result instead of each reducing every tile's value again (the same
// canonical functions, so the same bits).
__global__ __launch_bounds__(kB) void k_weight_gmax(
    const double *pmax, uint32_t nb, double *gmax)
{
    __shared__ double sd[8];
    const double m = blocks_max(pmax, nb, sd);
    if (threadIdx.x == 0)
        gmax[0] = m;
}

__global__ __launch_bounds__(kB) void k_weight_gsum(const double *ps,
    const double *ps2, uint32_t nb, const double *lmax, double *gsum,
    double *stats)
{
    __shared__ double sd[8];
    double a, b;
    tiles_sum2(ps, ps2, (int)nb, threadIdx.x, sd, a, b);
    if (threadIdx.x == 0) {
        gsum[0] = a;
        gsum[1] = b;
        stats[0] = __ddiv_rn(__dmul_rn(a, a), b);
        stats[1] = __dadd_rn(*lmax, log(a));
    }
}

// w = exp(w - lmax), lmax the max of the block maxima (taken by every
// workgroup, or from gmax; block 0 keeps it in *lmax); block partial sums of
// w and w^2
__global__ __launch_bounds__(kB) void k_weight_exp(uint64_t n, double *w,
    const double *pmax, uint32_t nb, const double *gmax, double *lmax,
    double *ps, double *ps2)
{
    __shared__ double sd[8];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    const double m = gmax ? gmax[0] : blocks_max(pmax, nb, sd);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *lmax = m;
    double a = 0.0, b = 0.0;
    double x[kItems];
    ld8(w, i0, n, x, 0.0);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < n) {
            const double e = exp(__dsub_rn(x[k], m));
            x[k] = e;
            a = __dadd_rn(a, e);
            b = __dadd_rn(b, __dmul_rn(e, e));
        }
    }
    st8(w, i0, n, x);
    group_sum2(a, b, threadIdx.x, sd);
    if (threadIdx.x == 0) {
        ps[blockIdx.x] = a;
        ps2[blockIdx.x] = b;
    }
}

// accw and the sum of squares from the block partials (the same sums in
// every workgroup), w *= 1/accw on the workgroup's tile; block 0 writes
// stats = {ESS, lmax + log(accw)}
__global__ __launch_bounds__(kB) void k_weight_scale(uint64_t n, double *w,
    const double *ps, const double *ps2, uint32_t nb, const double *gsum,
    const double *lmax, double *stats)
{
    __shared__ double sd[8];
    double a, b;
    if (gsum) { // k_weight_gsum wrote the sums and stats
        a = gsum[0];
        b = gsum[1];
    } else {
        tiles_sum2(ps, ps2, (int)nb, threadIdx.x, sd, a, b);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            stats[0] = __ddiv_rn(__dmul_rn(a, a), b);
            stats[1] = __dadd_rn(*lmax, log(a));
        }
    }
    const double sc = __ddiv_rn(1.0, a);
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    double x[kItems];
    ld8(w, i0, n, x, 0.0);
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        x[k] = __dmul_rn(sc, x[k]);
    st8(w, i0, n, x);
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
// term of element i with value a: a (mode 0, weights) or log(a) / (N - i)
// (mode 1, sorted uniforms)
__device__ __forceinline__ double scan_term(int mode, double a, uint64_t i,
    uint64_t N)
{
    return mode == 0 ? a : __ddiv_rn(log(a), (double)(N - i));
}

__global__ __launch_bounds__(kB) void k_dd_tiles(int mode, uint64_t n,
    uint64_t N, const double *a, double *th, double *tl)
{
    __shared__ double sd[8];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    dd loc{0.0, 0.0};
    double v[kItems];
    ld8(a, i0, n, v, 1.0);
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < n)
            loc = dd_add(loc, dd{scan_term(mode, v[k], i, N), 0.0});
    }
    dd tot;
    (void)group_excl_scan_dd(loc, threadIdx.x, sd, tot);
    if (threadIdx.x == 0) {
        th[blockIdx.x] = tot.hi;
        tl[blockIdx.x] = tot.lo;
    }
}

// Inclusive prefix of element i, rounded: the last element of a tile takes
// the next tile's offset (the same convention as the SMC samplers).  The
// tile's offsets come from the nt tile totals th/tl (tiles_offset_dd: the
// operations of a scan over the tiles, taken by every workgroup for
// itself); off_hi: this tile's offset rounded (the cumulative weight of the
// previous tile's last element).
// Many tiles: one workgroup writes every tile's offsets (tiles_scan_dd,
// tiles_scan_u2: the operations each workgroup would otherwise repeat over
// all tiles), and the tile kernels read their own.
constexpr uint32_t kTwoLevelTiles = 1024;

__global__ __launch_bounds__(kB) void k_tiles_scan_dd(const double *th,
    const double *tl, uint32_t nt, double *oh, double *ol)
{
    __shared__ double sd[16];
    tiles_scan_dd(true, th, tl, (int)nt, oh, ol, threadIdx.x, sd);
}

__global__ __launch_bounds__(kB) void k_tiles_scan_u2(const uint32_t *te,
    const uint32_t *tz, uint32_t nt, uint32_t *teo, uint32_t *tzo)
{
    __shared__ uint32_t su[8];
    tiles_scan_u2(true, te, tz, (int)nt, teo, tzo, threadIdx.x, su);
}

__device__ __forceinline__ void tile_prefix(int mode, uint64_t n, uint64_t N,
    const double *a, const double *th, const double *tl, uint32_t nt,
    const double *oh, const double *ol, double *P, double &off_hi)
{
    __shared__ double sd[16];
    const uint32_t tile = blockIdx.x;
    dd toff, tnext;
    if (oh) { // the offsets k_tiles_scan_dd wrote (the same operations)
        toff = dd{oh[tile], ol[tile]};
        tnext = dd{oh[tile + 1], ol[tile + 1]};
    } else {
        tiles_offset_dd(th, tl, (int)nt, (int)tile, (int)tile + 1,
            threadIdx.x, sd, toff, tnext);
    }
    off_hi = toff.hi;
    const uint64_t i0 = (uint64_t)tile * kTile + threadIdx.x * kItems;
    const uint64_t tend = (uint64_t)(tile + 1) * kTile < n
        ? (uint64_t)(tile + 1) * kTile
        : n;
    double t[kItems];
    ld8(a, i0, n, t, 1.0);
    dd loc{0.0, 0.0};
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        t[k] = i < n ? scan_term(mode, t[k], i, N) : 0.0;
        loc = dd_add(loc, dd{t[k], 0.0});
    }
    dd tot;
    const dd off = group_excl_scan_dd(loc, threadIdx.x, sd, tot);
    const dd base = dd_add(toff, off);
    dd run{0.0, 0.0};
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        run = dd_add(run, dd{t[k], 0.0});
        P[k] = i == tend - 1 ? tnext.hi : dd_add(base, run).hi;
    }
}

// sorted uniforms: r_j = 1 - exp(lm_j), lm = prefix of log(c_j)/(N-j)
// (resample.hpp:63-81)
__global__ __launch_bounds__(kB) void k_sorted_finish(uint64_t n,
    const double *c, const double *th, const double *tl, uint32_t nt,
    const double *oh, const double *ol, double *r)
{
    double P[kItems];
    double off_hi;
    tile_prefix(1, n, n, c, th, tl, nt, oh, ol, P, off_hi);
    __syncthreads();
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    double v[kItems];
    for (int k = 0; k < kItems; ++k)
        v[k] = __dsub_rn(1.0, exp(P[k]));
    st8(r, i0, n, v);
}

// #{j < M : u_j < S} for non-decreasing u, searched from a guess g in
// [0, M]: steps of 1, 2, 4, ... away from g until the partition point is
// bracketed, then bisection of the bracket.  The result is the partition
// point whatever g is; a guess d away from it costs about 2 log2(d) + 2
// loads instead of log2(M).
__device__ __forceinline__ uint64_t count_from(
    const double *u, uint64_t M, double S, uint64_t g)
{
    uint64_t lo, hi; // the partition point lies in [lo, hi]
    if (g < M && u[g] < S) {
        lo = g + 1;
        for (uint64_t step = 1;; step <<= 1) {
            const uint64_t j = lo + step - 1;
            if (j >= M) {
                hi = M;
                break;
            }
            if (!(u[j] < S)) {
                hi = j;
                break;
            }
            lo = j + 1;
        }
    } else {
        hi = g;
        for (uint64_t step = 1;; step <<= 1) {
            if (hi < step) {
                lo = 0;
                break;
            }
            const uint64_t j = hi - step;
            if (u[j] < S) {
                lo = j + 1;
                break;
            }
            hi = j;
        }
    }
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (u[mid] < S)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// floor(x) clamped to [0, m] (NaN -> 0)
__device__ __forceinline__ uint64_t clamp_floor(double x, uint64_t m)
{
    const double f = floor(x);
    return f > 0.0 ? (f < (double)m ? (uint64_t)f : m) : 0;
}

// ---- rep -> index, general M (resample.hpp:416-466) ----
// extras e_i = max(rep_i - 1, 0) for i < K, rep_i for i >= K (K = min(N,M));
// demands: zero slots d < K in order, then slots K..M-1.
__device__ __forceinline__ uint32_t extras(uint32_t rep, uint64_t i, uint64_t K)
{
    return i < K ? (rep > 1 ? rep - 1 : 0u) : rep;
}

// Replication numbers rep_i = cnt_i - cnt_{i-1} with cnt_i = #{u < S_i}
// for i < N-1 and cnt_{N-1} = M (resample.hpp:374-407).  For each of a
// thread's eight cumulative weights the uniforms g-1, g, g+1 around the
// guess g = floor(M S_i) are read first, all 24 loads in flight: a count
// the window brackets is final (always so for stratified and systematic
// uniforms away from rounding ties, where u_j lies in [j/M, (j+1)/M]).
// The other counts are searched in order (count_from), each from the
// previous count advanced by M (S_i - S_{i-1}): near the answer for sorted
// uniforms too.  The count before a thread's first particle is the
// previous thread's last, or for a tile's first thread the count at the
// previous tile's last particle, whose cumulative weight is this tile's
// offset.  With te != nullptr the tile's extra-copy and zero-slot totals
// of the rep -> index transform (K = min(N, M)) are written as well.
__global__ __launch_bounds__(kB) void k_counts(uint64_t N, uint64_t M,
    const double *w, const double *u, const double *th, const double *tl,
    uint32_t nt, const double *oh, const double *ol, uint32_t *rep,
    uint64_t K, uint32_t *te, uint32_t *tz)
{
    __shared__ uint32_t su[8];
    __shared__ uint32_t bnd[2];
    double P[kItems];
    double off_hi;
    tile_prefix(0, N, N, w, th, tl, nt, oh, ol, P, off_hi);
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    const double Md = (double)M;
    const double inf = __builtin_huge_val();
    uint64_t g[kItems];
    double ua[kItems], ub[kItems], uc[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const bool live = i0 + k < N - 1;
        g[k] = clamp_floor(__dmul_rn(P[k], Md), M);
        ua[k] = live && g[k] >= 1 ? u[g[k] - 1] : -inf;
        ub[k] = live && g[k] < M ? u[g[k]] : inf;
        uc[k] = live && g[k] + 1 < M ? u[g[k] + 1] : inf;
    }
    uint32_t c[kItems];
    uint64_t cc = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        c[k] = (uint32_t)M;
        if (i < N - 1) {
            const double S = P[k];
            if (ua[k] < S && !(ub[k] < S)) {
                cc = g[k];
            } else if (ub[k] < S && !(uc[k] < S)) {
                cc = g[k] + 1;
            } else {
                const uint64_t h = k == 0
                    ? g[0]
                    : cc + clamp_floor(
                               __dmul_rn(__dsub_rn(S, P[k - 1]), Md), M - cc);
                cc = count_from(u, M, S, h);
            }
            c[k] = (uint32_t)cc;
        }
    }
    // the counts at the tile's two boundaries (the previous tile's last
    // count and this tile's last), then the counts made non-decreasing
    // inside them, so no replication number is negative whatever the
    // rounding of the prefixes (smc_index.hpp)
    const uint64_t tend = (uint64_t)(blockIdx.x + 1) * kTile < N
        ? (uint64_t)(blockIdx.x + 1) * kTile
        : N;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (i0 + k == tend - 1)
            bnd[1] = c[k];
    if (threadIdx.x == 0) {
        const double S = off_hi;
        bnd[0] = blockIdx.x == 0 ? 0u
            : (uint32_t)count_from(u, M, S, clamp_floor(__dmul_rn(S, Md), M));
    }
    __syncthreads();
    const uint32_t clo = bnd[0], chi = bnd[1];
    uint32_t cmax = clo;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < N) {
            const uint32_t v = clamp_count(c[k], clo, chi);
            cmax = cmax > v ? cmax : v;
            c[k] = cmax;
        }
    }
    const uint32_t before = group_excl_max_u(cmax, clo, threadIdx.x, su);
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (i0 + k < N && c[k] < before)
            c[k] = before;
    uint32_t prev = before;
    uint32_t r[kItems];
    uint32_t es = 0, zs = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        r[k] = rep_from_counts(c[k], prev);
        prev = c[k];
        if (i < N) {
            es += extras(r[k], i, K);
            zs += (i < K && r[k] == 0) ? 1u : 0u;
        }
    }
    st8(rep, i0, N, r);
    if (te) {
        uint32_t a, b2;
        (void)group_excl_scan_u(es, threadIdx.x, su, a);
        (void)group_excl_scan_u(zs, threadIdx.x, su + 4, b2);
        if (threadIdx.x == 0) {
            te[blockIdx.x] = a;
            tz[blockIdx.x] = b2;
        }
    }
}

__global__ void k_rep_small(uint64_t N, uint64_t M, uint32_t *rep)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (uint64_t)gridDim.x * blockDim.x)
        rep[i] = (N == 1 && i == 0) ? (uint32_t)M : 0u;
}

__global__ __launch_bounds__(kB) void k_ri_tiles(uint64_t N, uint64_t K,
    const uint32_t *rep, uint32_t *te, uint32_t *tz)
{
    __shared__ uint32_t su[8];
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    uint32_t es = 0, zs = 0;
    uint32_t r[kItems];
    ld8(rep, i0, N, r, 0u);
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < N) {
            es += extras(r[k], i, K);
            zs += (i < K && r[k] == 0) ? 1u : 0u;
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

// sums of a[0..j) and b[0..j) and of all n entries of b (every thread)
__device__ __forceinline__ void tiles_before_u2(const uint32_t *a,
    const uint32_t *b, uint32_t n, uint32_t j, uint32_t *su, uint32_t &ab,
    uint32_t &bb, uint32_t &btot)
{
    uint32_t x = 0, y = 0, z = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kB) {
        x += i < j ? a[i] : 0u;
        y += i < j ? b[i] : 0u;
        z += b[i];
    }
    (void)group_excl_scan_u(x, threadIdx.x, su, ab);
    (void)group_excl_scan_u(y, threadIdx.x, su + 4, bb);
    (void)group_excl_scan_u(z, threadIdx.x, su + 8, btot);
}

// Inclusive extra-copy prefix E_i; idx_i = i for every slot i < K (the
// zero slots' entries are overwritten by k_ri_deal); zpos[z] = i for the
// zero slot i of rank z.
__global__ __launch_bounds__(kB) void k_ri_write(uint64_t N, uint64_t K,
    const uint32_t *rep, const uint32_t *te, const uint32_t *tz, uint32_t nt,
    const uint32_t *teo, const uint32_t *tzo, uint32_t *Eincl, uint32_t *idx,
    uint32_t *zpos)
{
    __shared__ uint32_t su[12];
    uint32_t eoff, zoff, ztot;
    if (teo) { // the offsets k_tiles_scan_u2 wrote
        eoff = teo[blockIdx.x];
        zoff = tzo[blockIdx.x];
        ztot = tzo[nt];
    } else {
        tiles_before_u2(te, tz, nt, blockIdx.x, su, eoff, zoff, ztot);
    }
    const uint64_t i0 = (uint64_t)blockIdx.x * kTile + threadIdx.x * kItems;
    uint32_t es = 0, zs = 0;
    uint32_t r[kItems];
    ld8(rep, i0, N, r, 0u);
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < N) {
            es += extras(r[k], i, K);
            zs += (i < K && r[k] == 0) ? 1u : 0u;
        }
    }
    uint32_t a, b;
    uint32_t ex = group_excl_scan_u(es, threadIdx.x, su, a) + eoff;
    uint32_t zx = group_excl_scan_u(zs, threadIdx.x, su + 4, b) + zoff;
    uint32_t E[kItems], own[kItems];
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        ex += i < N ? extras(r[k], i, K) : 0u;
        E[k] = ex;
        own[k] = (uint32_t)i;
        if (i < K && r[k] == 0) {
            zpos[zx] = (uint32_t)i;
            ++zx;
        }
    }
    st8(Eincl, i0, N, E);
    st8(idx, i0, K, own);
}

// The extra copies in ascending source order take the demands in order:
// extra copy r goes to the zero slot of rank r (r < ztot) or to slot
// K + r - ztot.  One workgroup per tile of sources: the tile's E in LDS, its
// ranks [E_{t0-1}, E_{t0+n-1}) dealt round-robin to the threads, each
// rank's source (the first j with E_j > r) from LDS: when the tile holds at
// most kTile ranks, every source marks its first rank and a max-scan over
// the ranks fills in the rest (sources are in rank order); otherwise by
// bisection.  Slots past M (a replication vector that does not sum to M)
// are not written.
__global__ __launch_bounds__(kB) void k_ri_deal(uint64_t N, uint64_t K,
    uint64_t M, const uint32_t *Eincl, const uint32_t *zpos,
    const uint32_t *tz, uint32_t nt, const uint32_t *tzo, uint32_t *idx)
{
    __shared__ uint32_t Es[kTile];
    __shared__ uint32_t Src[kTile];
    __shared__ uint32_t su[4];
    uint32_t ztot;
    if (tzo) {
        ztot = tzo[nt];
    } else {
        uint32_t z = 0;
        for (uint32_t i = threadIdx.x; i < nt; i += kB)
            z += tz[i];
        (void)group_excl_scan_u(z, threadIdx.x, su, ztot);
    }
    const uint64_t t0 = (uint64_t)blockIdx.x * kTile;
    const uint32_t n = (uint32_t)(N - t0 < (uint64_t)kTile ? N - t0 : kTile);
    for (uint32_t q = threadIdx.x; q < n; q += kB)
        Es[q] = Eincl[t0 + q];
    const uint32_t beg = t0 == 0 ? 0u : Eincl[t0 - 1];
    __syncthreads();
    // ranks past ztot + (M - K) have no destination: a vector that is not a
    // replication vector (extras beyond M) ends its loop there
    const uint32_t end = rank_end(Es[n - 1], (uint64_t)ztot + (M - K));
    if (end > beg && end - beg <= (uint32_t)kTile) { // uniform over the group
        // marks: source q's first rank E_{q-1} (q's extras are the ranks
        // [E_{q-1}, E_q)); rank 0 of the tile is always marked
        const uint32_t q0 = threadIdx.x * kItems;
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            Src[q0 + k] = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint32_t q = q0 + k;
            if (q < n) {
                const uint32_t first = q == 0 ? beg : Es[q - 1];
                if (Es[q] > first && first < end)
                    Src[first - beg] = q;
            }
        }
        __syncthreads();
        uint32_t v[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            v[k] = Src[q0 + k];
#pragma unroll
        for (int k = 1; k < kItems; ++k)
            v[k] = v[k] > v[k - 1] ? v[k] : v[k - 1];
        const uint32_t before =
            group_excl_max_u(v[kItems - 1], 0u, threadIdx.x, su);
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            Src[q0 + k] = v[k] > before ? v[k] : before;
        __syncthreads();
        for (uint32_t r = beg + threadIdx.x; r < end; r += kB) {
            const uint64_t dest =
                r < ztot ? (uint64_t)zpos[r] : K + (uint64_t)(r - ztot);
            if (dest < M)
                idx[dest] = (uint32_t)(t0 + Src[r - beg]);
        }
        return;
    }
    for (uint32_t r = beg + threadIdx.x; r < end; r += kB) {
        const uint32_t lo = rank_source(Es, n, r); // Es[n - 1] >= end > r
        const uint64_t dest =
            r < ztot ? (uint64_t)zpos[r] : K + (uint64_t)(r - ztot);
        if (dest < M)
            idx[dest] = (uint32_t)(t0 + lo);
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

// column-major: column c of the new state (M rows) is column c of the old
// one (N rows) gathered by idx
template <typename T>
__global__ void k_select_cols(uint64_t M, uint64_t N, uint64_t dim,
    const uint32_t *idx, const T *src, T *dst)
{
    const uint64_t total = M * dim;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
         q < total; q += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = q / M, j = q % M;
        dst[q] = src[c * N + idx[j]];
    }
}

uint32_t grid_for(uint64_t n, uint32_t cap = 8192)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(n, 256), cap));
}

// The library's own stream-ordered pool on each device, created on first
// use and kept until the process ends.  Its release threshold is unbounded:
// memory freed by a call stays in the pool for the next one, also across a
// stream synchronisation (the residual schemes synchronise inside every call
// to read R), instead of going back to the driver and being mapped again.
// The pool holds at most the largest scratch one call has needed.  The
// process's default pool is left as the caller set it.
constexpr int kMaxPoolDevices = 64;
std::mutex g_pool_mu;
hipMemPool_t g_pool[kMaxPoolDevices] = {};

int scratch_pool(hipMemPool_t &pool)
{
    int dev = 0;
    MCKL_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxPoolDevices)
        return set_error(MCKL_HIP_EINVAL, "device ordinal beyond 64");
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (!g_pool[dev]) {
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t p = nullptr;
        MCKL_HIP_TRY(hipMemPoolCreate(&p, &props));
        uint64_t keep = UINT64_MAX;
        const hipError_t err =
            hipMemPoolSetAttribute(p, hipMemPoolAttrReleaseThreshold, &keep);
        if (err != hipSuccess) {
            (void)hipMemPoolDestroy(p);
            MCKL_HIP_TRY(err);
        }
        g_pool[dev] = p;
    }
    pool = g_pool[dev];
    return MCKL_HIP_OK;
}

// RAII for stream-ordered scratch from the library's pool
struct Scratch {
    hipStream_t st;
    void *p = nullptr;
    explicit Scratch(hipStream_t s) : st(s) {}
    int alloc(size_t bytes)
    {
        hipMemPool_t pool = nullptr;
        const int rc = scratch_pool(pool);
        if (rc)
            return rc;
        MCKL_HIP_TRY(hipMallocFromPoolAsync(
            &p, std::max<size_t>(bytes, 256), pool, st));
        return MCKL_HIP_OK;
    }
    ~Scratch()
    {
        if (p)
            (void)hipFreeAsync(p, st);
    }
};

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

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
    // past 1024 tiles every workgroup reducing every tile's values again
    // costs more than two one-workgroup launches
    const bool two_level = nt > 1024;
    double *gmax = two_level ? lmax + 1 : nullptr;
    double *gsum = two_level ? lmax + 2 : nullptr;
    k_weight_pre<<<nt, kB, 0, st>>>(mode, n, v, w, pmax);
    if (two_level)
        k_weight_gmax<<<1, kB, 0, st>>>(pmax, nt, gmax);
    k_weight_exp<<<nt, kB, 0, st>>>(n, w, pmax, nt, gmax, lmax, ps, ps2);
    if (two_level)
        k_weight_gsum<<<1, kB, 0, st>>>(ps, ps2, nt, lmax, gsum, stats);
    k_weight_scale<<<nt, kB, 0, st>>>(n, w, ps, ps2, nt, gsum, lmax, stats);
    MCKL_HIP_CHECK_LAUNCH("weight kernels");
    return MCKL_HIP_OK;
}

// double-double tile totals of the scan terms (nt of them); the consumers
// take their tile offsets from these
int dd_tiles(int mode, size_t n, const double *a, double *th, double *tl,
    hipStream_t st)
{
    const uint32_t nt = (uint32_t)ceil_div(n, kTile);
    k_dd_tiles<<<nt, kB, 0, st>>>(mode, n, n, a, th, tl);
    MCKL_HIP_CHECK_LAUNCH("k_dd_tiles");
    return MCKL_HIP_OK;
}

// Device workspace of one resampling call, carved from one allocation:
// double-double tile totals (the weights' and the sorted uniforms' scans),
// the rep -> index tile totals, the extra-copy prefix and zero-slot
// positions, the systematic draw and, for sorted uniforms, a copy of the
// draws.
struct RsWork {
    double *th, *tl;
    uint32_t *te, *tz;
    uint32_t *E, *zpos;
    double *u01;
    double *c;
    double *oh, *ol;     // tile offsets (many tiles: k_tiles_scan_dd)
    uint32_t *teo, *tzo; // tile offsets (many tiles: k_tiles_scan_u2)
};

constexpr size_t rs_work_bytes(size_t N, size_t M, bool sorted)
{
    const size_t nt = std::max(ceil_div(N, kTile), ceil_div(M, kTile)) + 1;
    return 4 * align256(nt * sizeof(double)) +
        4 * align256(nt * sizeof(uint32_t)) +
        2 * align256(std::max<size_t>(N, 1) * sizeof(uint32_t)) + 256 +
        (sorted ? align256(std::max<size_t>(M, 1) * sizeof(double)) : 0);
}

void rs_work_carve(void *p, size_t N, size_t M, bool sorted, RsWork &w)
{
    const size_t nt = std::max(ceil_div(N, kTile), ceil_div(M, kTile)) + 1;
    const size_t od = align256(nt * sizeof(double));
    const size_t ou = align256(nt * sizeof(uint32_t));
    const size_t on = align256(std::max<size_t>(N, 1) * sizeof(uint32_t));
    char *b = static_cast<char *>(p);
    w.th = reinterpret_cast<double *>(b);
    w.tl = reinterpret_cast<double *>(b + od);
    b += 2 * od;
    w.te = reinterpret_cast<uint32_t *>(b);
    w.tz = reinterpret_cast<uint32_t *>(b + ou);
    b += 2 * ou;
    w.oh = reinterpret_cast<double *>(b);
    w.ol = reinterpret_cast<double *>(b + od);
    b += 2 * od;
    w.teo = reinterpret_cast<uint32_t *>(b);
    w.tzo = reinterpret_cast<uint32_t *>(b + ou);
    b += 2 * ou;
    w.E = reinterpret_cast<uint32_t *>(b);
    w.zpos = reinterpret_cast<uint32_t *>(b + on);
    b += 2 * on;
    w.u01 = reinterpret_cast<double *>(b);
    b += 256;
    w.c = sorted ? reinterpret_cast<double *>(b) : nullptr;
}

// The helpers below take a workspace from the caller or allocate their own.
struct OwnWork {
    Scratch s;
    RsWork w;
    explicit OwnWork(hipStream_t st) : s(st) {}
    int get(const RsWork *given, size_t N, size_t M, bool sorted,
        const RsWork *&out)
    {
        if (given) {
            out = given;
            return MCKL_HIP_OK;
        }
        const int rc = s.alloc(rs_work_bytes(N, M, sorted));
        if (rc)
            return rc;
        rs_work_carve(s.p, N, M, sorted, w);
        out = &w;
        return MCKL_HIP_OK;
    }
};

int sequence(int scheme, mckl_hip_engine *e, size_t n, double *r,
    hipStream_t st, const RsWork *given = nullptr)
{
    const bool sorted = scheme == MCKL_HIP_MULTINOMIAL;
    if (!sorted && scheme != MCKL_HIP_SYSTEMATIC) {
        if (n == 0)
            return MCKL_HIP_OK;
        const int rc = rand_canonical_f64(e, n, r, st);
        if (rc)
            return rc;
        k_stratified<<<grid_for(n), 256, 0, st>>>(n, r);
        MCKL_HIP_CHECK_LAUNCH("k_stratified");
        return MCKL_HIP_OK;
    }
    if (sorted && n == 0)
        return MCKL_HIP_OK;
    OwnWork own(st);
    const RsWork *w = nullptr;
    int rc = own.get(given, sorted ? n : 0, n, sorted, w);
    if (rc)
        return rc;
    if (!sorted) {
        // one U01Canonical<double> draw (two words), resample.hpp:258-268
        rc = rand_canonical_f64(e, 1, w->u01, st);
        if (rc)
            return rc;
        if (n == 0)
            return MCKL_HIP_OK;
        k_systematic<<<grid_for(n), 256, 0, st>>>(n, w->u01, r);
        MCKL_HIP_CHECK_LAUNCH("k_systematic");
        return MCKL_HIP_OK;
    }
    // the draws go straight to the workspace (k_sorted_finish reads them
    // there and writes r)
    rc = rand_canonical_f64(e, n, w->c, st);
    if (rc)
        return rc;
    const uint32_t nt = (uint32_t)ceil_div(n, kTile);
    rc = dd_tiles(1, n, w->c, w->th, w->tl, st);
    if (rc)
        return rc;
    const bool two = nt > kTwoLevelTiles;
    if (two)
        k_tiles_scan_dd<<<1, kB, 0, st>>>(w->th, w->tl, nt, w->oh, w->ol);
    k_sorted_finish<<<nt, kB, 0, st>>>(n, w->c, w->th, w->tl, nt,
        two ? w->oh : nullptr, two ? w->ol : nullptr, r);
    MCKL_HIP_CHECK_LAUNCH("k_sorted_finish");
    return MCKL_HIP_OK;
}

// rep from the weights and M ascending uniforms; with tiles, the
// rep -> index tile totals go to w->te/w->tz as well (returns in *tiles
// whether they were written)
int trans_u01_rep(size_t N, size_t M, const double *wt, const double *u,
    uint32_t *rep, hipStream_t st, const RsWork *given = nullptr,
    bool tiles = false, bool *tiles_done = nullptr)
{
    if (tiles_done)
        *tiles_done = false;
    if (N == 0)
        return MCKL_HIP_OK;
    if (N == 1 || M == 0) {
        // resample.hpp:387-397: rep = {M} for N == 1, zeros for M == 0
        k_rep_small<<<grid_for(N), 256, 0, st>>>(N, M, rep);
        MCKL_HIP_CHECK_LAUNCH("k_rep_small");
        return MCKL_HIP_OK;
    }
    OwnWork own(st);
    const RsWork *w = nullptr;
    int rc = own.get(given, N, M, false, w);
    if (rc)
        return rc;
    const uint32_t nt = (uint32_t)ceil_div(N, kTile);
    rc = dd_tiles(0, N, wt, w->th, w->tl, st);
    if (rc)
        return rc;
    const bool two = nt > kTwoLevelTiles;
    if (two)
        k_tiles_scan_dd<<<1, kB, 0, st>>>(w->th, w->tl, nt, w->oh, w->ol);
    k_counts<<<nt, kB, 0, st>>>(N, M, wt, u, w->th, w->tl, nt,
        two ? w->oh : nullptr, two ? w->ol : nullptr, rep, std::min(N, M),
        tiles ? w->te : nullptr, tiles ? w->tz : nullptr);
    MCKL_HIP_CHECK_LAUNCH("k_counts");
    if (tiles_done)
        *tiles_done = tiles;
    return MCKL_HIP_OK;
}

// stable rep -> index; have_tiles: w->te/w->tz already hold the tile totals
int rep_index(size_t N, size_t M, const uint32_t *rep, uint32_t *idx,
    hipStream_t st, const RsWork *given = nullptr, bool have_tiles = false)
{
    if (N == 0 || M == 0)
        return MCKL_HIP_OK;
    const size_t K = std::min(N, M);
    const uint32_t nt = (uint32_t)ceil_div(N, kTile);
    OwnWork own(st);
    const RsWork *w = nullptr;
    int rc = own.get(given, N, M, false, w);
    if (rc)
        return rc;
    if (!have_tiles)
        k_ri_tiles<<<nt, kB, 0, st>>>(N, K, rep, w->te, w->tz);
    const bool two = nt > kTwoLevelTiles;
    if (two)
        k_tiles_scan_u2<<<1, kB, 0, st>>>(w->te, w->tz, nt, w->teo, w->tzo);
    k_ri_write<<<nt, kB, 0, st>>>(N, K, rep, w->te, w->tz, nt,
        two ? w->teo : nullptr, two ? w->tzo : nullptr, w->E, idx, w->zpos);
    k_ri_deal<<<nt, kB, 0, st>>>(N, K, M, w->E, w->zpos, w->tz, nt,
        two ? w->tzo : nullptr, idx);
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
    double x[kItems];
    uint32_t ig[kItems];
    ld8(w, i0, N, x, 0.0);
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        double ip = 0.0, r = 0.0;
        if (i < N) {
            r = modf(__dmul_rn(Md, x[k]), &ip);
            loc = dd_add(loc, dd{r, 0.0});
            isum += (unsigned long long)ip;
        }
        x[k] = r;
        ig[k] = (uint32_t)ip;
    }
    st8(resid, i0, N, x);
    st8(integ, i0, N, ig);
    dd tot;
    (void)group_excl_scan_dd(loc, threadIdx.x, sd, tot);
    isum = wave_sum_u(isum);
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
    isum = wave_sum_u(isum);
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
// ws: a workspace of at least residual_ws_bytes(N, M, base) bytes, whose
// RsWork part the caller may reuse for rep -> index afterwards.
constexpr size_t residual_ws_bytes(size_t N, size_t M, int base)
{
    const size_t nt = ceil_div(N, kTile) + 1;
    return 3 * align256(nt * sizeof(double)) +
        align256(std::max<size_t>(N, 1) * sizeof(double)) +
        align256(std::max<size_t>(N, 1) * sizeof(uint32_t)) + 512 +
        align256(std::max<size_t>(M, 1) * sizeof(double)) +
        rs_work_bytes(N, M, base == MCKL_HIP_MULTINOMIAL);
}

int residual_rep(int base, size_t N, size_t M, mckl_hip_engine *e,
    const double *w, uint32_t *rep, hipStream_t st, void *ws, RsWork &work)
{
    const uint32_t nt = (uint32_t)ceil_div(N, kTile);
    const size_t o = align256((nt + 1) * sizeof(double));
    const size_t on = align256(std::max<size_t>(N, 1) * sizeof(double));
    const size_t oi = align256(std::max<size_t>(N, 1) * sizeof(uint32_t));
    const size_t om = align256(std::max<size_t>(M, 1) * sizeof(double));
    char *base_p = static_cast<char *>(ws);
    double *th = reinterpret_cast<double *>(base_p);
    double *tl = reinterpret_cast<double *>(base_p + o);
    unsigned long long *ti =
        reinterpret_cast<unsigned long long *>(base_p + 2 * o);
    double *resid = reinterpret_cast<double *>(base_p + 3 * o);
    uint32_t *integ = reinterpret_cast<uint32_t *>(base_p + 3 * o + on);
    double *mul = reinterpret_cast<double *>(base_p + 3 * o + on + oi);
    unsigned long long *dR = reinterpret_cast<unsigned long long *>(
        base_p + 3 * o + on + oi + 256);
    double *u = reinterpret_cast<double *>(base_p + 3 * o + on + oi + 512);
    rs_work_carve(base_p + 3 * o + on + oi + 512 + om, N, M,
        base == MCKL_HIP_MULTINOMIAL, work);
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
    int rc = sequence(base, e, R, u, st, &work);
    if (rc)
        return rc;
    rc = trans_u01_rep(N, R, resid, u, rep, st, &work);
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

template <typename T>
int select_cols_impl(size_t M, size_t N, size_t dim, const uint32_t *idx,
    const T *src, T *dst, hipStream_t st)
{
    if (M == 0 || dim == 0)
        return MCKL_HIP_OK;
    if (!idx || !src || !dst)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (src == dst)
        return set_error(MCKL_HIP_EINVAL, "select is out of place");
    if (N == 0 || N > UINT32_MAX)
        return set_error(MCKL_HIP_EINVAL, "select: N out of range");
    if (dim > ((uint64_t)1 << 40) / std::max(M, N))
        return set_error(MCKL_HIP_EINVAL, "select: size out of range");
    k_select_cols<T><<<grid_for(M * dim), 256, 0, st>>>(M, N, dim, idx, src,
        dst);
    MCKL_HIP_CHECK_LAUNCH("k_select_cols");
    return MCKL_HIP_OK;
}

// Workspace bytes of one ResampleAlgorithm call: the M uniforms and the
// helpers' arrays (plain schemes), or the residual form's arrays.
constexpr size_t resample_ws_bytes(int scheme, size_t N, size_t M)
{
    return scheme >= MCKL_HIP_RESIDUAL
        ? residual_ws_bytes(N, M, scheme - MCKL_HIP_RESIDUAL)
        : align256(std::max<size_t>(M, 1) * sizeof(double)) +
            rs_work_bytes(N, M, scheme == MCKL_HIP_MULTINOMIAL);
}

int resample_check(int scheme, size_t N, size_t M, const mckl_hip_engine *e,
    const double *w, const uint32_t *rep)
{
    const bool residual = scheme >= MCKL_HIP_RESIDUAL &&
        scheme <= MCKL_HIP_RESIDUAL_SYSTEMATIC;
    const int rc = check_scheme(residual ? scheme - MCKL_HIP_RESIDUAL : scheme);
    if (rc)
        return rc;
    if (!e || !rep)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    // N == 1 and M == 0 never read the weights (rep = {M} or zeros)
    if (!w && N > 1 && M > 0)
        return set_error(MCKL_HIP_EINVAL, "NULL weights");
    if (N >= 0x7FFFFFFFu || M >= 0x7FFFFFFFu)
        return set_error(MCKL_HIP_EINVAL, "sizes must be < 2^31");
    return MCKL_HIP_OK;
}

// ResampleAlgorithm::operator() + rep -> index on a workspace of
// resample_ws_bytes(scheme, N, M) bytes; ws == NULL: allocate one on the
// stream.  Every kernel of the call writes its part of the workspace before a
// later kernel of the same call reads it, so nothing needs to start at zero.
int resample_impl(int scheme, size_t N, size_t M, mckl_hip_engine *e,
    const double *d_w, uint32_t *d_rep, uint32_t *d_idx, hipStream_t st,
    void *ws)
{
    Scratch s(st);
    const bool residual = scheme >= MCKL_HIP_RESIDUAL;
    if (residual && N == 0)
        return MCKL_HIP_OK;
    if (!ws) {
        const int rc = s.alloc(resample_ws_bytes(scheme, N, M));
        if (rc)
            return rc;
        ws = s.p;
    }
    int rc;
    if (residual) {
        RsWork w;
        rc = residual_rep(scheme - MCKL_HIP_RESIDUAL, N, M, e, d_w, d_rep, st,
            ws, w);
        if (rc)
            return rc;
        if (d_idx)
            rc = rep_index(N, M, d_rep, d_idx, st, &w);
        return rc;
    }
    // one workspace for the uniforms and every helper
    const bool sorted = scheme == MCKL_HIP_MULTINOMIAL;
    const size_t ou = align256(std::max<size_t>(M, 1) * sizeof(double));
    double *u = static_cast<double *>(ws);
    RsWork w;
    rs_work_carve(static_cast<char *>(ws) + ou, N, M, sorted, w);
    rc = sequence(scheme, e, M, u, st, &w);
    if (rc)
        return rc;
    bool tiles = false;
    rc = trans_u01_rep(N, M, d_w, u, d_rep, st, &w, d_idx != nullptr, &tiles);
    if (rc)
        return rc;
    if (d_idx)
        rc = rep_index(N, M, d_rep, d_idx, st, &w, tiles);
    return rc;
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

// ---- resampling workspace for callers that cannot allocate ----
// The workspace of mckl_hip_resample_noalloc lives in this array, sized for
// N, M <= kNoAllocMax under every scheme (about 36 MiB).  Each device holds
// its own copy until the process ends.  In host code `scratch` names a host
// stand-in, so its device address is looked up once per device
// (noalloc_workspace) and only that address is handed to kernels.
constexpr size_t kNoAllocMax = size_t(1) << 20;
constexpr size_t noalloc_bytes()
{
    size_t b = 0;
    for (int s = MCKL_HIP_MULTINOMIAL; s <= MCKL_HIP_RESIDUAL_SYSTEMATIC; ++s)
        b = std::max(b, resample_ws_bytes(s, kNoAllocMax, kNoAllocMax));
    return b;
}
constexpr size_t kNoAllocBytes = noalloc_bytes();

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices];

// the current device's address of `scratch`, looked up on first use
int noalloc_workspace(void **out)
{
    int dev = -1;
    MCKL_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(MCKL_HIP_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        MCKL_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return set_error(MCKL_HIP_EHIP, "no device address for scratch");
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
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
    const int rc = resample_check(scheme, N, M, e, d_w, d_rep);
    if (rc)
        return rc;
    return resample_impl(scheme, N, M, e, d_w, d_rep, d_idx,
        as_stream(stream), nullptr);
}

int mckl_hip_resample_noalloc(int scheme, size_t N, size_t M,
    mckl_hip_engine *e, const double *d_w, uint32_t *d_rep, uint32_t *d_idx,
    void *stream)
{
    int rc = resample_check(scheme, N, M, e, d_w, d_rep);
    if (rc)
        return rc;
    if (N > kNoAllocMax || M > kNoAllocMax)
        return set_error(MCKL_HIP_EINVAL,
            "resample_noalloc: N = %zu, M = %zu exceed the workspace limit "
            "%zu", N, M, kNoAllocMax);
    void *ws = nullptr;
    rc = noalloc_workspace(&ws);
    if (rc)
        return rc;
    return resample_impl(scheme, N, M, e, d_w, d_rep, d_idx,
        as_stream(stream), ws);
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

int mckl_hip_select_cols_f32(size_t M, size_t N, size_t dim,
    const uint32_t *d_idx, const float *d_src, float *d_dst, void *stream)
{
    return select_cols_impl(M, N, dim, d_idx, d_src, d_dst, as_stream(stream));
}

int mckl_hip_select_cols_f64(size_t M, size_t N, size_t dim,
    const uint32_t *d_idx, const double *d_src, double *d_dst, void *stream)
{
    return select_cols_impl(M, N, dim, d_idx, d_src, d_dst, as_stream(stream));
}

} // extern "C"
