// smc.hip -- SMC samplers on gfx950.
//
//  * plain sampler: one workgroup per 2048-particle tile, a handful of
//    launches per step, no host synchronisation (the ESS test is taken on the
//    device and the resampling kernels exit when it is negative), so the
//    whole T loop can be captured into one HIP graph;
//  * batched sampler: thousands of independent filters of uneven sizes in one
//    persistent launch, one workgroup per filter taken from a device queue
//    ordered by decreasing particle count;
// both built from the per-tile stages of smc_tiles.hpp, so a filter gives
// the same bits in either sampler.
#include "common.hpp"
#include "smc_tiles.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

namespace mckl_hip {

// Device state of the plain sampler.  Per-step values live in two slots
// indexed by step parity: step t reads slot (t-1)&1 and writes slot t&1, so
// the workgroups of one launch that read the previous values never race the
// one that writes the new ones.
struct SmcState {
    double logz;
    double S;         // the step's weight total rounded once (residual steps)
    double C[2];      // normaliser of the carried weights for the next step
    float m[2];       // max log weight of the step
    int resampled[2]; // the step resampled (applies to the next move)
    int err;          // a look-back wait gave up (results are invalid)
};

// -------------------------------------------------------------------------
// plain sampler kernels (grid = ntiles workgroups of 256 threads).  No
// device-wide fences: a value every tile needs (max, sums, the resampling
// decision, tile offsets) is recomputed by each workgroup from the per-tile
// arrays with the canonical cross-tile functions, and only workgroup 0
// writes it to the state; the one cross-tile prefix that depends on the
// step's counts (the tiles' extra-copy and zero-slot totals) is a
// decoupled look-back on relaxed atomic words.
// -------------------------------------------------------------------------
struct RsBuf {
    double *tdh, *tdl; // tile dd totals of a (unscaled)
    uint32_t *te, *tz; // tile extra-copy / zero-slot counts
    uint32_t *cnt;     // cumulative counts (multinomial, residual)
    double *tmh, *tml; // multinomial: tile dd totals of the log spacings
    double *u;         // multinomial: the step's sorted uniforms
    double *uh, *ul;   // multinomial: scaled tile dd totals r T
    uint32_t *ti;      // residual: tile sums of the integral parts
    double *trh, *trl; // residual: tile dd totals of the residuals
    uint32_t *src;     // source of each extra-copy rank (tile_index_store)
    uint64_t *lbw;     // look-back words, 2 x ntiles (by step parity)
};

__global__ __launch_bounds__(kGroup) void k_init(
    SmcArgs A, float *x, float *lw, SmcState *st, RsBuf B)
{
    tile_init(A, blockIdx.x, x, lw, threadIdx.x);
    if (threadIdx.x == 0) {
        B.lbw[blockIdx.x] = 0;
        B.lbw[A.ntiles + blockIdx.x] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->logz = 0.0;
        st->C[0] = (double)A.N;
        st->m[0] = 0.0f;
        st->resampled[0] = 0;
        st->err = 0;
    }
}

// move + weight of a tile, then the tile's sums of a_i = exp_f(lw_i - m_k)
// about its own max m_k (the weights' scale r_k needs the max over tiles and
// is applied by the next kernel)
__global__ __launch_bounds__(kGroup) void k_move(SmcArgs A, uint32_t t,
    const float *xin, float *xout, float *lw, const uint32_t *idx,
    const SmcState *st, float *tmax, double *ts, double *ts2, RsBuf B)
{
    __shared__ __align__(16) unsigned char lds[kScratch];
    const GroupLds L = group_lds(lds);
    const int p = (t - 1) & 1;
    float l[kItems];
    const float mt = tile_move<true>(A, t, blockIdx.x, xin, xout, lw, idx,
        B.src, st->resampled[p], st->m[p], threadIdx.x, L.f, l);
    double S, S2;
    dd et;
    tile_sums_l(A, blockIdx.x, l, mt, threadIdx.x, L.d, S, S2, et);
    if (threadIdx.x == 0) {
        tmax[blockIdx.x] = mt;
        ts[blockIdx.x] = S;
        ts2[blockIdx.x] = S2;
        B.tdh[blockIdx.x] = et.hi;
        B.tdl[blockIdx.x] = et.lo;
    }
}

__device__ __forceinline__ float tiles_max(
    const float *tmax, uint32_t nt, int gt, float *ldsf)
{
    float m = -__builtin_huge_valf();
    for (uint32_t i = gt; i < nt; i += kGroup)
        m = fmaxf(m, tmax[i]);
    return group_max(m, gt, ldsf);
}

// Every workgroup reduces the tile sums and takes the step's decision;
// workgroup 0 records it.  When the step resamples, each workgroup derives
// its tile's prefix offsets and counts the uniforms below its cumulative
// weights.  All global loads the workgroup needs are issued up front (its
// tile's log weights, its chunk of every per-tile array), so the kernel
// waits on memory once; the reductions are those of tiles_max, tiles_sum2
// and tiles_offset_dd, operation for operation.
__host__ __device__ constexpr bool is_residual(int scheme)
{
    return scheme == MCKL_HIP_RESIDUAL ||
        scheme == MCKL_HIP_RESIDUAL_STRATIFIED ||
        scheme == MCKL_HIP_RESIDUAL_SYSTEMATIC;
}

// Decoupled look-back (single-pass prefix over the tiles of one launch):
// tile k publishes its (extra-copy, zero-slot) totals as an aggregate, sums
// its predecessors' words back to the nearest inclusive prefix, then
// publishes its own inclusive prefix.  A word packs a status (bit 62
// aggregate, bit 63 inclusive) with two 31-bit totals, so one relaxed
// agent-scope atomic load or store carries value and status together and no
// fence is needed.  Workgroups start in index order and every one of a
// launch is resident (one per tile, a few per CU), so a predecessor always
// makes progress; the wait is bounded all the same and, if it gives up,
// sets st->err and counts the missing tiles as empty (every index the
// kernels then derive still stays below N).
constexpr uint64_t kLbAgg = 1ull << 62;
constexpr uint64_t kLbIncl = 1ull << 63;
constexpr uint32_t kLbSpinLimit = 1u << 20;

__device__ __forceinline__ uint64_t lb_pack(
    uint64_t status, uint32_t e, uint32_t z)
{
    return status | ((uint64_t)(e & 0x7FFFFFFFu) << 31) |
        (uint64_t)(z & 0x7FFFFFFFu);
}

__device__ __forceinline__ uint64_t lb_load(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void lb_store(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// tile k's aggregate (or, for tile 0, inclusive) word; one thread calls it
__device__ __forceinline__ void lb_publish(
    uint64_t *w, uint32_t k, uint32_t te, uint32_t tz)
{
    lb_store(&w[k], lb_pack(k == 0 ? kLbIncl : kLbAgg, te, tz));
}

// exclusive prefix (eoff, zoff) of tile k's totals (te, tz) over the tiles
// before it, after lb_publish.  Each round reads kLbQ predecessor words per
// thread (a window
// of kLbQ x 256 tiles, one memory round trip), waits until each holds a
// status, and adds the words up to the nearest inclusive one; a launch of at
// most 1024 tiles finishes in one round.  Every thread of the workgroup
// calls it.
constexpr int kLbQ = 4;

__device__ __forceinline__ void tile_lookback(uint64_t *w, uint32_t k,
    uint32_t te, uint32_t tz, int gt, uint32_t *ldsu, int *err,
    uint32_t &eoff, uint32_t &zoff)
{
    uint32_t es = 0, zs = 0, spins = 0;
    bool failed = false;
    int64_t top = (int64_t)k - 1; // the window's nearest predecessor
    while (top >= 0) {
        uint64_t v[kLbQ];
#pragma unroll
        for (int q = 0; q < kLbQ; ++q) {
            const int64_t j = top - (gt + q * kGroup);
            v[q] = j >= 0 ? lb_load(&w[j]) : kLbIncl; // before tile 0: zero
        }
#pragma unroll
        for (int q = 0; q < kLbQ; ++q) {
            const int64_t j = top - (gt + q * kGroup);
            while (!(v[q] & (kLbAgg | kLbIncl))) {
                if (++spins > kLbSpinLimit) {
                    failed = true;
                    v[q] = kLbIncl;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
                v[q] = lb_load(&w[j]);
            }
        }
        // distance of the nearest inclusive word; words up to it count
        uint32_t dmin = 0xFFFFFFFFu;
#pragma unroll
        for (int q = 0; q < kLbQ; ++q) {
            const uint32_t d = (uint32_t)(gt + q * kGroup);
            if ((v[q] & kLbIncl) && d < dmin)
                dmin = d;
        }
        dmin = group_min_u(dmin, gt, ldsu + 8);
#pragma unroll
        for (int q = 0; q < kLbQ; ++q) {
            if ((uint32_t)(gt + q * kGroup) <= dmin) {
                es += (uint32_t)((v[q] >> 31) & 0x7FFFFFFFu);
                zs += (uint32_t)(v[q] & 0x7FFFFFFFu);
            }
        }
        if (dmin != 0xFFFFFFFFu)
            break;
        top -= (int64_t)kLbQ * kGroup;
    }
    if (failed)
        *err = 1;
    group_sum2_u(es, zs, gt, ldsu);
    if (gt == 0 && k != 0)
        lb_store(&w[k], lb_pack(kLbIncl, es + te, zs + tz));
    eoff = es;
    zoff = zs;
}

// one instantiation per scheme S (a compile-time copy of A.scheme), so each
// launch carries only its own scheme's code
template <int S>
__global__ __launch_bounds__(kGroup) void k_rs_counts(SmcArgs A, uint32_t t,
    const float *lw, const float *tmax, const double *ts, const double *ts2,
    RsBuf B, SmcState *st, double *ess_hist, int32_t *res_hist,
    uint32_t *idx)
{
    A.scheme = S;
    __shared__ __align__(16) unsigned char lds[kScratch];
    __shared__ __align__(16) uint32_t Es[kTile];
    __shared__ uint32_t Srcl[kTile];
    const GroupLds L = group_lds(lds);
    const int gt = threadIdx.x;
    const int nt = (int)A.ntiles;
    const uint32_t k = blockIdx.x;
    const int p = (t - 1) & 1, c = t & 1;
    // this tile's look-back word of the next step (last read by the
    // previous step's launch, which has finished)
    if (gt == 0)
        lb_store(&B.lbw[(size_t)(c ^ 1) * nt + k], 0);
    float l[kItems];
    load8(lw, tile_first(k, gt), A.N, l, 0.0f);
    const double Cprev = st->C[p];
    const float tk = tmax[k];
    const int chunk = (nt + kGroup - 1) / kGroup;
    const int b = gt * chunk;
    const int e = b + chunk < nt ? b + chunk : nt;
    // up to 1024 tiles (chunks of at most kPre per thread) the thread's
    // tile values are loaded before the group max, in one round trip with
    // the tile maxima; the arithmetic below is the same either way
    constexpr int kPre = 4;
    const bool pre = chunk <= kPre;
    float pt[kPre];
    double ps[kPre], ps2[kPre], ph[kPre], pl[kPre];
    float mx = -__builtin_huge_valf();
    if (pre) {
#pragma unroll
        for (int q = 0; q < kPre; ++q) {
            const int i = b + q;
            if (i < e) {
                pt[q] = tmax[i];
                ps[q] = ts[i];
                ps2[q] = ts2[i];
                ph[q] = B.tdh[i];
                pl[q] = B.tdl[i];
                mx = fmaxf(mx, pt[q]);
            }
        }
    } else {
        for (int i = gt; i < nt; i += kGroup)
            mx = fmaxf(mx, tmax[i]);
    }
    const float m = group_max(mx, gt, L.f);
    // sums of the scaled tile sums: r_i S_i and r_i^2 S2_i (the decision,
    // log Z and the next normaliser), r_i T_i in double-double (the offsets
    // and, rounded once, the total that normalises the cumulative weights)
    double sa = 0.0, sb = 0.0;
    dd loc{0.0, 0.0};
    dd pu[kPre]; // the scaled dd tile totals r_i T_i (pre path)
    if (pre) {
#pragma unroll
        for (int q = 0; q < kPre; ++q) {
            pu[q] = dd{0.0, 0.0};
            if (b + q < e) {
                const double r = tile_scale(pt[q], m);
                sa = __dadd_rn(sa, __dmul_rn(r, ps[q]));
                sb = __dadd_rn(sb, __dmul_rn(__dmul_rn(r, r), ps2[q]));
                pu[q] = dd_mul_d(dd{ph[q], pl[q]}, r);
                loc = dd_add(loc, pu[q]);
            }
        }
    } else {
        for (int i = b; i < e; ++i) {
            const double r = tile_scale(tmax[i], m);
            sa = __dadd_rn(sa, __dmul_rn(r, ts[i]));
            sb = __dadd_rn(sb, __dmul_rn(__dmul_rn(r, r), ts2[i]));
            loc = dd_add(loc, dd_mul_d(dd{B.tdh[i], B.tdl[i]}, r));
        }
    }
    group_sum2(sa, sb, gt, L.d);
    const double rs = sa, rs2 = sb;
    // the decision in every workgroup; the log Z increment (two logs) only
    // where it is recorded
    StepOut o;
    o.ess = __ddiv_rn(__dmul_rn(rs, rs), rs2);
    o.resample = o.ess < A.tauN ? 1 : 0;
    if (blockIdx.x == 0 && gt == 0) {
        o = step_finalize(A, m, rs, rs2, Cprev);
        st->logz = __dadd_rn(st->logz, o.logz_inc);
        st->m[c] = m;
        st->resampled[c] = o.resample;
        st->C[c] = o.resample ? (double)A.N : rs;
        if (ess_hist)
            ess_hist[t - 1] = o.ess;
        if (res_hist)
            res_hist[t - 1] = o.resample;
    }
    if (!o.resample)
        return;
    const double rk = tile_scale(tk, m);
    if constexpr (S == MCKL_HIP_MULTINOMIAL) {
        // multinomial: tile totals of the log spacings; the uniforms and
        // the counts follow in k_rs_uniforms and k_rs_mcounts, which read
        // this tile's scaled total from B.uh/B.ul
        dd mt;
        tile_mterms(A, k, t, gt, L.d, mt);
        if (gt == 0) {
            B.tmh[k] = mt.hi;
            B.tml[k] = mt.lo;
            const dd u = dd_mul_d(dd{B.tdh[k], B.tdl[k]}, rk);
            B.uh[k] = u.hi;
            B.ul[k] = u.lo;
        }
        return;
    }
    // offsets of tiles k and k + 1 (tiles_offset_dd on the chunk sums)
    dd total;
    const dd off = group_excl_scan_dd(loc, gt, L.d, total);
    double *x = L.d + 8;
    {
        dd run = off;
        if (pre) {
#pragma unroll
            for (int q = 0; q < kPre; ++q) {
                const int i = b + q;
                if (i < e) {
                    if (i == (int)k) {
                        x[0] = run.hi;
                        x[1] = run.lo;
                    }
                    if (i == (int)k + 1)
                        x[2] = run.hi;
                    run = dd_add(run, pu[q]);
                }
            }
        } else {
            for (int i = b; i < e; ++i) {
                if (i == (int)k) {
                    x[0] = run.hi;
                    x[1] = run.lo;
                }
                if (i == (int)k + 1)
                    x[2] = run.hi;
                run = dd_add(run, dd_mul_d(dd{B.tdh[i], B.tdl[i]},
                                      tile_scale(tmax[i], m)));
            }
        }
        if (gt == 0 && (int)k + 1 == nt)
            x[2] = total.hi;
    }
    __syncthreads();
    const dd toff{x[0], x[1]};
    const double next_hi = x[2];
    __syncthreads();
    const double invS = __ddiv_rn(1.0, total.hi);
    if constexpr (is_residual(S)) {
        // residual: the tile's integral parts and residual total; the
        // counts follow in k_rs_rcounts (after k_rs_rterms and
        // k_rs_uniforms for residual multinomial)
        uint32_t isum;
        dd rt;
        tile_residual_sums(A, k, l, tk, rk, invS, gt, L.d, L.u, isum, rt);
        if (gt == 0) {
            B.ti[k] = isum;
            B.trh[k] = rt.hi;
            B.trl[k] = rt.lo;
            if (k == 0)
                st->S = total.hi;
        }
        return;
    }
    // systematic / stratified: the counts, the tile's place among the
    // extra copies and zero slots by look-back, and the rep -> index
    // transform (parents as codes, resolved by the next move); the tile's
    // own part of the transform runs while its predecessors publish
    const UniformSeq U = make_uniform_seq(A, t);
    uint32_t es, zs, cs[kItems], prev;
    float ta[kItems];
    dd ttoff;
    tile_counts_local(A, k, l, tk, gt, L.d, ta, ttoff);
    tile_counts_c(A, k, ta, ttoff, rk, invS, toff, next_hi, U, gt, L.u, cs,
        prev, es, zs);
    uint64_t *w = B.lbw + (size_t)c * nt;
    if (gt == 0)
        lb_publish(w, k, es, zs);
    TileIndex ti;
    tile_index_local(A, k, cs, prev, gt, L.u, Es, Srcl, ti);
    uint32_t eoff, zoff;
    tile_lookback(w, k, es, zs, gt, L.u, &st->err, eoff, zoff);
    tile_index_store(A, k, ti, eoff, zoff, idx, B.src, gt, Es, Srcl);
}

// R = M - sum of the tile sums of the integral parts (all threads)
__device__ __forceinline__ uint32_t residual_R(
    const SmcArgs &A, const uint32_t *ti, int gt, uint32_t *ldsu)
{
    uint32_t is = 0, zero = 0;
    for (uint32_t i = gt; i < A.ntiles; i += kGroup)
        is += ti[i];
    group_sum2_u(is, zero, gt, ldsu);
    return is <= A.N ? A.N - is : 0u;
}

// residual multinomial: tile totals of the log spacings of the R uniforms
__global__ __launch_bounds__(kGroup) void k_rs_rterms(SmcArgs A, uint32_t t,
    const SmcState *st, RsBuf B)
{
    if (!st->resampled[t & 1])
        return;
    __shared__ __align__(16) unsigned char lds[kScratch];
    const GroupLds L = group_lds(lds);
    const uint32_t k = blockIdx.x;
    const uint32_t R = residual_R(A, B.ti, threadIdx.x, L.u);
    dd mt;
    tile_mterms(A, k, t, threadIdx.x, L.d, mt, R);
    if (threadIdx.x == 0) {
        B.tmh[k] = mt.hi;
        B.tml[k] = mt.lo;
    }
}

// multinomial: each tile's sorted uniforms from the prefix of log spacings
__global__ __launch_bounds__(kGroup) void k_rs_uniforms(SmcArgs A, uint32_t t,
    const SmcState *st, RsBuf B)
{
    if (!st->resampled[t & 1])
        return;
    __shared__ __align__(16) unsigned char lds[kScratch];
    const GroupLds L = group_lds(lds);
    const uint32_t k = blockIdx.x;
    const uint32_t M = is_residual(A.scheme)
        ? residual_R(A, B.ti, threadIdx.x, L.u)
        : A.N;
    dd off, next;
    tiles_offset_dd(B.tmh, B.tml, (int)A.ntiles, (int)k, (int)k + 1,
        threadIdx.x, L.d, off, next);
    tile_muniforms(A, k, t, off, B.u, threadIdx.x, L.d, M);
}

// multinomial: counts of the sorted uniforms below each cumulative weight
__global__ __launch_bounds__(kGroup) void k_rs_mcounts(SmcArgs A, uint32_t t,
    const float *lw, const float *tmax, const SmcState *st, RsBuf B)
{
    const int c = t & 1;
    if (!st->resampled[c])
        return;
    __shared__ __align__(16) unsigned char lds[kScratch];
    const GroupLds L = group_lds(lds);
    const int gt = threadIdx.x;
    const int nt = (int)A.ntiles;
    const float m = st->m[c];
    const double rs = tiles_total_dd(B.uh, B.ul, nt, gt, L.d).hi;
    const uint32_t k = blockIdx.x;
    dd off, next;
    tiles_offset_dd(B.uh, B.ul, nt, (int)k, (int)k + 1, gt, L.d, off, next);
    const double invS = __ddiv_rn(1.0, rs);
    const UniformSeq U = make_uniform_seq(A, t, B.u);
    uint32_t es, zs;
    tile_counts(A, k, lw, tmax[k], tile_scale(tmax[k], m), invS, off, next.hi,
        U, B.cnt, gt, L.d, L.u, es, zs);
    if (gt == 0) {
        B.te[k] = es;
        B.tz[k] = zs;
    }
}

// residual steps: R = M - sum of the integral parts, the residual total and
// the tile's offsets (tiles_offset_dd's operations on the chunk sums), then
// the cumulative counts of the tile (tile_rcounts)
__global__ __launch_bounds__(kGroup) void k_rs_rcounts(SmcArgs A, uint32_t t,
    const float *lw, const float *tmax, const SmcState *st, RsBuf B)
{
    const int c = t & 1;
    if (!st->resampled[c])
        return;
    __shared__ __align__(16) unsigned char lds[kScratch];
    const GroupLds L = group_lds(lds);
    const int gt = threadIdx.x;
    const int nt = (int)A.ntiles;
    const uint32_t k = blockIdx.x;
    float l[kItems];
    load8(lw, tile_first(k, gt), A.N, l, 0.0f);
    const float m = st->m[c];
    const double invS = __ddiv_rn(1.0, st->S);
    const int chunk = (nt + kGroup - 1) / kGroup;
    const int b = gt * chunk;
    const int e = b + chunk < nt ? b + chunk : nt;
    uint32_t isum = 0, ioff = 0;
    dd loc{0.0, 0.0};
    for (int i = b; i < e; ++i) {
        isum += B.ti[i];
        ioff += i < (int)k ? B.ti[i] : 0u;
        loc = dd_add(loc, dd{B.trh[i], B.trl[i]});
    }
    group_sum2_u(isum, ioff, gt, L.u);
    const uint32_t R = isum <= A.N ? A.N - isum : 0u;
    dd total;
    const dd off = group_excl_scan_dd(loc, gt, L.d, total);
    double *x = L.d + 8;
    {
        dd run = off;
        for (int i = b; i < e; ++i) {
            if (i == (int)k) {
                x[0] = run.hi;
                x[1] = run.lo;
            }
            if (i == (int)k + 1)
                x[2] = run.hi;
            run = dd_add(run, dd{B.trh[i], B.trl[i]});
        }
        if (gt == 0 && (int)k + 1 == nt)
            x[2] = total.hi;
    }
    __syncthreads();
    const dd roff{x[0], x[1]};
    const double rnext_hi = x[2];
    __syncthreads();
    UniformSeq U;
    U.u = A.scheme == MCKL_HIP_RESIDUAL ? B.u : nullptr;
    U.scheme = A.scheme == MCKL_HIP_RESIDUAL_SYSTEMATIC ? MCKL_HIP_SYSTEMATIC
        : A.scheme == MCKL_HIP_RESIDUAL_STRATIFIED      ? MCKL_HIP_STRATIFIED
                                                        : MCKL_HIP_MULTINOMIAL;
    U.k0 = A.k0;
    U.k1 = A.k1;
    U.t = t;
    U.M = R;
    U.delta = R ? __ddiv_rn(1.0, (double)R) : 0.0;
    U.u0d = U.scheme == MCKL_HIP_SYSTEMATIC
        ? __dmul_rn(smc_uniform(A.k0, A.k1, t, 0), U.delta)
        : 0.0;
    const double invR = R ? __ddiv_rn(1.0, total.hi) : 0.0;
    const float mk = tmax[k];
    uint32_t es, zs;
    tile_rcounts(A, k, l, mk, tile_scale(mk, m), invS, ioff, roff, rnext_hi,
        invR, U, B.cnt, gt, L.d, L.u, es, zs);
    if (gt == 0) {
        B.te[k] = es;
        B.tz[k] = zs;
    }
}

// exclusive prefix of a u32 tile array up to tile k (all threads)
__device__ __forceinline__ void tiles_prefix_u2(const uint32_t *a,
    const uint32_t *b, uint32_t k, int gt, uint32_t *ldsu, uint32_t &pa,
    uint32_t &pb)
{
    uint32_t sa = 0, sb = 0;
    for (uint32_t i = gt; i < k; i += kGroup) {
        sa += a[i];
        sb += b[i];
    }
    group_sum2_u(sa, sb, gt, ldsu);
    pa = sa;
    pb = sb;
}

// rep -> index of the multinomial and residual schemes, whose counts come
// from kernels of their own (the tile offsets from the tile totals)
__global__ __launch_bounds__(kGroup) void k_rs_index(SmcArgs A, uint32_t t,
    const SmcState *st, RsBuf B, uint32_t *idx)
{
    if (!st->resampled[t & 1])
        return;
    __shared__ __align__(16) unsigned char lds[kScratch];
    __shared__ __align__(16) uint32_t Es[kTile];
    __shared__ uint32_t Srcl[kTile];
    const GroupLds L = group_lds(lds);
    uint32_t c[kItems], prev;
    load_counts(A, blockIdx.x, B.cnt, threadIdx.x, c, prev);
    uint32_t eoff, zoff;
    tiles_prefix_u2(B.te, B.tz, blockIdx.x, threadIdx.x, L.u, eoff, zoff);
    TileIndex ti;
    tile_index_local(A, blockIdx.x, c, prev, threadIdx.x, L.u, Es, Srcl, ti);
    tile_index_store(A, blockIdx.x, ti, eoff, zoff, idx, B.src, threadIdx.x,
        Es, Srcl);
}

// -------------------------------------------------------------------------
// batched sampler: one workgroup (G groups of 256 threads) per filter
// -------------------------------------------------------------------------
constexpr int kMaxTiles = 64; // N <= 131072 per filter in the batched form

struct FilterDesc {
    SmcArgs A;
    uint64_t off; // particle offset into the shared buffers
};

struct BatchLds {
    float tmax[kMaxTiles];
    double r[kMaxTiles]; // tile scales exp_f(m_k - m)
    double ts[kMaxTiles], ts2[kMaxTiles];
    double tdh[kMaxTiles], tdl[kMaxTiles];
    double toh[kMaxTiles + 1], tol[kMaxTiles + 1];
    uint32_t te[kMaxTiles], tz[kMaxTiles];
    uint32_t teo[kMaxTiles + 1], tzo[kMaxTiles + 1];
    double tmh[kMaxTiles], tml[kMaxTiles];
    uint32_t ti[kMaxTiles];              // residual: integral parts
    double trh[kMaxTiles], trl[kMaxTiles]; // residual: residual totals
    int flag;
    int filter;
};

// Sorted uniforms of step t for one filter (M of them: N, or R of a
// residual step): tile totals of the log spacings, their offsets, then the
// uniforms (the plain sampler's k_rs_counts/k_rs_rterms, k_rs_uniforms).
// Leaves S.toh/S.tol overwritten.  Every thread of the workgroup calls it.
template <int G>
__device__ __forceinline__ void batch_muniforms(const SmcArgs &A, uint32_t t,
    int nt, int rounds, int grp, int gt, const GroupLds &L, BatchLds &S,
    double *u, uint32_t M)
{
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        dd mt;
        tile_mterms(A, (uint32_t)(k < nt ? k : nt), t, gt, L.d, mt, M);
        if (gt == 0 && k < nt) {
            S.tmh[k] = mt.hi;
            S.tml[k] = mt.lo;
        }
        __syncthreads();
    }
    tiles_scan_dd(grp == 0, S.tmh, S.tml, nt, S.toh, S.tol, gt, L.d);
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        const int kk = k < nt ? k : nt;
        tile_muniforms(
            A, (uint32_t)kk, t, dd{S.toh[kk], S.tol[kk]}, u, gt, L.d, M);
    }
    __syncthreads();
}

// Counts of a residual step for one filter, the operations of the plain
// sampler's k_rs_counts (residual branch), k_rs_rterms, k_rs_uniforms and
// k_rs_rcounts: per tile the integral parts and residual totals, R, the R
// sorted uniforms for residual multinomial, the residual offsets, then the
// tile's cumulative counts.  Leaves S.te/S.tz set and S.toh/S.tol holding the
// residual offsets.  Every thread of the workgroup calls it.
template <int G>
__device__ __forceinline__ void batch_rcounts(const SmcArgs &A, uint32_t t,
    int nt, int rounds, int grp, int gt, const GroupLds &L, BatchLds &S,
    const float *lw, double *u, uint32_t *cnt, double invS)
{
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        const int kk = k < nt ? k : nt;
        float l[kItems];
        load8(lw, tile_first((uint32_t)kk, gt), A.N, l, 0.0f);
        uint32_t isum;
        dd rt;
        tile_residual_sums(A, (uint32_t)kk, l, kk < nt ? S.tmax[kk] : 0.0f,
            kk < nt ? S.r[kk] : 0.0, invS, gt, L.d, L.u, isum, rt);
        if (gt == 0 && k < nt) {
            S.ti[k] = isum;
            S.trh[k] = rt.hi;
            S.trl[k] = rt.lo;
        }
        __syncthreads();
    }
    uint32_t isum = 0;
    for (int i = 0; i < nt; ++i)
        isum += S.ti[i];
    const uint32_t R = isum <= A.N ? A.N - isum : 0u;
    if (A.scheme == MCKL_HIP_RESIDUAL)
        batch_muniforms<G>(A, t, nt, rounds, grp, gt, L, S, u, R);
    tiles_scan_dd(grp == 0, S.trh, S.trl, nt, S.toh, S.tol, gt, L.d);
    UniformSeq U;
    U.u = A.scheme == MCKL_HIP_RESIDUAL ? u : nullptr;
    U.scheme = A.scheme == MCKL_HIP_RESIDUAL_SYSTEMATIC ? MCKL_HIP_SYSTEMATIC
        : A.scheme == MCKL_HIP_RESIDUAL_STRATIFIED      ? MCKL_HIP_STRATIFIED
                                                        : MCKL_HIP_MULTINOMIAL;
    U.k0 = A.k0;
    U.k1 = A.k1;
    U.t = t;
    U.M = R;
    U.delta = R ? __ddiv_rn(1.0, (double)R) : 0.0;
    U.u0d = U.scheme == MCKL_HIP_SYSTEMATIC
        ? __dmul_rn(smc_uniform(A.k0, A.k1, t, 0), U.delta)
        : 0.0;
    const double invR = R ? __ddiv_rn(1.0, S.toh[nt]) : 0.0;
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        const int kk = k < nt ? k : nt;
        uint32_t ioff = 0;
        for (int i = 0; i < kk; ++i)
            ioff += S.ti[i];
        float l[kItems];
        load8(lw, tile_first((uint32_t)kk, gt), A.N, l, 0.0f);
        uint32_t es, zs;
        tile_rcounts(A, (uint32_t)kk, l, kk < nt ? S.tmax[kk] : 0.0f,
            kk < nt ? S.r[kk] : 0.0, invS, ioff, dd{S.toh[kk], S.tol[kk]},
            kk < nt ? S.toh[kk + 1] : 0.0, invR, U, cnt, gt, L.d, L.u, es, zs);
        if (gt == 0 && k < nt) {
            S.te[k] = es;
            S.tz[k] = zs;
        }
        __syncthreads();
    }
}

template <int G>
__global__ __launch_bounds__(G *kGroup, G == 1 ? 2 : 1) void k_batch(const FilterDesc *fd,
    const uint32_t *order, uint32_t nfilters, uint32_t *queue, float *xa,
    float *xb, float *lwb, uint32_t *idxb, uint32_t *cntb, uint32_t *Eb,
    uint32_t *zb, double *ub, double *logz_out)
{
    __shared__ __align__(16) unsigned char scratch[G * kScratch];
    __shared__ BatchLds S;
    __shared__ uint32_t Es[G * kTile]; // tile_scatter_ranks, one per group
    const int grp = threadIdx.x / kGroup;
    const int gt = threadIdx.x % kGroup;
    const GroupLds L = group_lds(scratch + grp * kScratch);
    for (;;) {
        if (threadIdx.x == 0) {
            const uint32_t q = atomicAdd(queue, 1u);
            S.filter = q < nfilters ? (int)order[q] : -1;
        }
        __syncthreads();
        const int f = S.filter;
        __syncthreads();
        if (f < 0)
            return;
        const SmcArgs A = fd[f].A;
        const uint64_t off = fd[f].off;
        float *x0 = xa + off, *x1 = xb + off, *lw = lwb + off;
        uint32_t *idx = idxb + off, *cnt = cntb + off, *Eincl = Eb + off;
        uint32_t *zpos = zb + off;
        const int nt = (int)A.ntiles;
        const int rounds = (nt + G - 1) / G;

        for (int k = grp; k < nt; k += G)
            tile_init(A, (uint32_t)k, x0, lw, gt);
        __syncthreads();
        double logz = 0.0, C = (double)A.N;
        float mprev = 0.0f;
        int resampled = 0;
        // Groups without a tile in a round take tile index nt (past the
        // end): every particle index is >= N, so they touch no memory but
        // still meet every barrier of the group functions.
        for (uint32_t t = 1; t <= A.T; ++t) {
            const float *xin = (t & 1) ? x0 : x1;
            float *xout = (t & 1) ? x1 : x0;
            // move + weight, with the tile sums about each tile's max
            for (int r = 0; r < rounds; ++r) {
                const int k = r * G + grp;
                const uint32_t kk = (uint32_t)(k < nt ? k : nt);
                float l[kItems];
                const float mt = tile_move<false>(A, t, kk, xin, xout, lw,
                    idx, nullptr, resampled, mprev, gt, L.f, l);
                double s, s2;
                dd et;
                tile_sums_l(A, kk, l, mt, gt, L.d, s, s2, et);
                if (gt == 0 && k < nt) {
                    S.tmax[k] = mt;
                    S.ts[k] = s;
                    S.ts2[k] = s2;
                    S.tdh[k] = et.hi;
                    S.tdl[k] = et.lo;
                }
                __syncthreads();
            }
            float m = -__builtin_huge_valf();
            for (int i = 0; i < nt; ++i)
                m = fmaxf(m, S.tmax[i]);
            // scale the tile sums by r_k = exp_f(m_k - m), as the plain
            // sampler's counting kernel does on the fly
            for (int i = threadIdx.x; i < nt; i += G * kGroup) {
                const double rr = tile_scale(S.tmax[i], m);
                S.r[i] = rr;
                S.ts[i] = __dmul_rn(rr, S.ts[i]);
                S.ts2[i] = __dmul_rn(__dmul_rn(rr, rr), S.ts2[i]);
                const dd u = dd_mul_d(dd{S.tdh[i], S.tdl[i]}, rr);
                S.tdh[i] = u.hi;
                S.tdl[i] = u.lo;
            }
            __syncthreads();
            double rs, rs2;
            tiles_sum2(S.ts, S.ts2, nt, gt, L.d, rs, rs2);
            const StepOut o = step_finalize(A, m, rs, rs2, C);
            logz = __dadd_rn(logz, o.logz_inc);
            C = o.resample ? (double)A.N : rs;
            mprev = m;
            resampled = o.resample;
            if (resampled) {
                double *u = ub ? ub + off : nullptr;
                if (A.scheme == MCKL_HIP_MULTINOMIAL) // sorted uniforms first
                    batch_muniforms<G>(A, t, nt, rounds, grp, gt, L, S, u, A.N);
                tiles_scan_dd(grp == 0, S.tdh, S.tdl, nt, S.toh, S.tol, gt, L.d);
                // the weight total rounded once normalises the cumulative
                // weights (the plain sampler's counting kernel does the same)
                const double invS = __ddiv_rn(1.0, S.toh[nt]);
                if (is_residual(A.scheme)) {
                    batch_rcounts<G>(A, t, nt, rounds, grp, gt, L, S, lw, u,
                        cnt, invS);
                } else {
                    const UniformSeq U = make_uniform_seq(A, t, u);
                    for (int r = 0; r < rounds; ++r) {
                        const int k = r * G + grp;
                        uint32_t es, zs;
                        const int kk = k < nt ? k : nt;
                        tile_counts(A, (uint32_t)kk, lw,
                            kk < nt ? S.tmax[kk] : 0.0f,
                            kk < nt ? S.r[kk] : 0.0, invS,
                            dd{S.toh[kk], S.tol[kk]},
                            kk < nt ? S.toh[kk + 1] : 0.0, U, cnt, gt, L.d,
                            L.u, es, zs);
                        if (gt == 0 && k < nt) {
                            S.te[k] = es;
                            S.tz[k] = zs;
                        }
                        __syncthreads();
                    }
                }
                tiles_scan_u2(grp == 0, S.te, S.tz, nt, S.teo, S.tzo, gt, L.u);
                for (int r = 0; r < rounds; ++r) {
                    const int k = r * G + grp;
                    const int kk = k < nt ? k : nt;
                    tile_index(A, (uint32_t)kk, cnt, S.teo[kk], S.tzo[kk],
                        Eincl, idx, zpos, gt, L.u);
                    __syncthreads();
                }
                for (int r = 0; r < rounds; ++r) {
                    const int k = r * G + grp;
                    tile_scatter_ranks(A, (uint32_t)(k < nt ? k : nt), Eincl,
                        zpos, idx, gt, Es + grp * kTile);
                }
            }
        }
        if (threadIdx.x == 0)
            logz_out[f] = logz;
        __syncthreads();
    }
}

// Selection estimator of the moments x^1..x^K (SMCEstimator at step 0,
// algorithm/smc.hpp:89-118 and :389-394): after the move and weighting of
// step t, per tile the double-double sums of e_i and of e_i x_i^j with the
// samplers' weights e_i = a_i r_k; products split exactly (two_prod).
// Partials: [(t-1) (K+1) + q][tile].
constexpr int kMaxMoments = 4;

__global__ __launch_bounds__(kGroup) void k_est_moments(SmcArgs A, uint32_t t,
    const float *x, const float *lw, const float *tmax, int K, double *eph,
    double *epl)
{
    __shared__ __align__(16) unsigned char lds[kScratch];
    const GroupLds L = group_lds(lds);
    const int gt = threadIdx.x;
    float l[kItems], xv[kItems];
    load8(lw, tile_first(blockIdx.x, gt), A.N, l, 0.0f);
    load8(x, tile_first(blockIdx.x, gt), A.N, xv, 0.0f);
    const float m = tiles_max(tmax, A.ntiles, gt, L.f);
    const float mk = tmax[blockIdx.x];
    const double rk = tile_scale(mk, m);
    dd acc[kMaxMoments + 1];
#pragma unroll
    for (int q = 0; q <= kMaxMoments; ++q)
        acc[q] = dd{0.0, 0.0};
    const uint64_t i0 = tile_first(blockIdx.x, gt);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < A.N) {
            const double e = __dmul_rn((double)smc_exp_f(__fsub_rn(l[k], mk)),
                rk); // exact: a product of two floats
            acc[0] = dd_add(acc[0], dd{e, 0.0});
            const double xd = (double)xv[k];
            double p = 1.0;
#pragma unroll
            for (int q = 1; q <= kMaxMoments; ++q) {
                if (q <= K) {
                    p = __dmul_rn(p, xd);
                    acc[q] = dd_add(acc[q], two_prod(e, p));
                }
            }
        }
    }
    double *sm = L.d; // 2 * (K+1) * 4 doubles
    const int wv = gt >> 6;
#pragma unroll
    for (int q = 0; q <= kMaxMoments; ++q) {
        if (q <= K) {
            const dd s = wave_sum_dd(acc[q]);
            if ((gt & 63) == 0) {
                sm[2 * (q * kWaves + wv)] = s.hi;
                sm[2 * (q * kWaves + wv) + 1] = s.lo;
            }
        }
    }
    __syncthreads();
    if (gt <= K) {
        dd s{0.0, 0.0};
        for (int w = 0; w < kWaves; ++w)
            s = dd_add(s, dd{sm[2 * (gt * kWaves + w)],
                             sm[2 * (gt * kWaves + w) + 1]});
        const uint64_t o =
            ((uint64_t)(t - 1) * (K + 1) + gt) * A.ntiles + blockIdx.x;
        eph[o] = s.hi;
        epl[o] = s.lo;
    }
}

// one workgroup per step: totals over tiles, est[t][j-1] = sum e x^j / sum e
__global__ __launch_bounds__(kGroup) void k_est_moments_final(uint32_t nt,
    int K, const double *eph, const double *epl, double *est)
{
    __shared__ double sm[2 * kWaves];
    const uint32_t t = blockIdx.x;
    double tot0 = 1.0;
    for (int q = 0; q <= K; ++q) {
        const uint64_t base = ((uint64_t)t * (K + 1) + q) * nt;
        dd s{0.0, 0.0};
        for (uint32_t i = threadIdx.x; i < nt; i += kGroup)
            s = dd_add(s, dd{eph[base + i], epl[base + i]});
        s = wave_sum_dd(s);
        if ((threadIdx.x & 63) == 0) {
            sm[2 * (threadIdx.x >> 6)] = s.hi;
            sm[2 * (threadIdx.x >> 6) + 1] = s.lo;
        }
        __syncthreads();
        dd g{0.0, 0.0};
        for (int w = 0; w < kWaves; ++w)
            g = dd_add(g, dd{sm[2 * w], sm[2 * w + 1]});
        __syncthreads();
        if (q == 0)
            tot0 = __dadd_rn(g.hi, g.lo);
        else if (threadIdx.x == 0)
            est[(uint64_t)t * K + q - 1] = dd_div_d(g, tot0);
    }
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        p[i] = 0;
}

// -------------------------------------------------------------------------
// host side
// -------------------------------------------------------------------------
static int model_const(int model, const double *th, ModelConst &mc)
{
    if (model != 1 && model != 2)
        return set_error(MCKL_HIP_EINVAL, "unknown model %d", model);
    if (!th)
        return set_error(MCKL_HIP_EINVAL, "theta is NULL");
    if (!(th[2] > 0) || !(th[1] > 0))
        return set_error(MCKL_HIP_EINVAL, "scale parameters must be > 0");
    if (model == 2 && !(std::fabs(th[0]) < 1))
        return set_error(MCKL_HIP_EINVAL, "SV requires |phi| < 1");
    const double l2pi = std::log(2.0 * 3.141592653589793);
    mc.model = model;
    mc.a = (float)th[0];
    mc.s = (float)th[1];
    mc.cst = (float)(-std::log(th[2]) - 0.5 * l2pi);
    if (model == 1) {
        mc.sd0 = 1.0f;
        mc.inv_sy = (float)(1.0 / th[2]);
    } else {
        mc.sd0 = (float)std::sqrt(th[1] * th[1] / (1.0 - th[0] * th[0]));
        mc.inv_sy = 0.0f;
    }
    return MCKL_HIP_OK;
}

static void obs_const(int model, const double *th, const double *y, size_t T,
    float *out)
{
    for (size_t t = 0; t < T; ++t)
        out[t] = model == 1 ? (float)y[t] : (float)(y[t] * y[t] / (th[2] * th[2]));
}

static int check_common(size_t N, size_t T, double threshold, int scheme)
{
    if (N == 0 || N >= (size_t)0x7FFFFFFF)
        return set_error(MCKL_HIP_EINVAL, "particle count %zu out of range", N);
    if (T == 0)
        return set_error(MCKL_HIP_EINVAL, "T must be > 0");
    if (scheme != MCKL_HIP_MULTINOMIAL && scheme != MCKL_HIP_STRATIFIED &&
        scheme != MCKL_HIP_SYSTEMATIC &&
        !is_residual(scheme))
        return set_error(MCKL_HIP_EINVAL,
            "resampling scheme %d is not provided by this sampler", scheme);
    if (std::isnan(threshold))
        return set_error(MCKL_HIP_EINVAL, "threshold is NaN");
    return MCKL_HIP_OK;
}

} // namespace mckl_hip

using namespace mckl_hip;

struct mckl_hip_smc {
    SmcArgs A;
    size_t T;
    float *x[2] = {nullptr, nullptr};
    float *lw = nullptr, *obs = nullptr, *tmax = nullptr;
    uint32_t *idx = nullptr;
    double *ts = nullptr, *ts2 = nullptr;
    RsBuf B;
    SmcState *st = nullptr;
    double *ess = nullptr;
    int32_t *res = nullptr;
    int kmom = 0;                                  // estimated moments
    double *eph = nullptr, *epl = nullptr, *est = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // recorded on the caller's stream after each run: result calls wait for
    // it, not for the whole device, and read through a non-blocking stream
    // of their own (the null stream would wait for every blocking stream)
    hipEvent_t done = nullptr;
    hipStream_t rs = nullptr;
};

static void smc_drop_graph(mckl_hip_smc *s)
{
    if (s->exec)
        (void)hipGraphExecDestroy(s->exec);
    if (s->graph)
        (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
}

static void smc_free(mckl_hip_smc *s)
{
    smc_drop_graph(s);
    if (s->done)
        (void)hipEventDestroy(s->done);
    if (s->rs)
        (void)hipStreamDestroy(s->rs);
    void *ptrs[] = {s->x[0], s->x[1], s->lw, s->obs, s->tmax, s->idx, s->ts,
        s->ts2, s->B.tdh, s->B.tdl, s->B.te, s->B.tz, s->B.cnt, s->B.tmh, s->B.tml, s->B.u, s->B.uh, s->B.ul, s->B.ti,
        s->B.trh, s->B.trl, s->B.src, s->B.lbw, s->st, s->ess, s->res, s->eph,
        s->epl, s->est};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    delete s;
}

template <typename... Args>
static void launch_rs_counts(int scheme, dim3 grid, hipStream_t st,
    Args... a)
{
    switch (scheme) {
    case MCKL_HIP_MULTINOMIAL:
        k_rs_counts<MCKL_HIP_MULTINOMIAL><<<grid, kGroup, 0, st>>>(a...);
        break;
    case MCKL_HIP_STRATIFIED:
        k_rs_counts<MCKL_HIP_STRATIFIED><<<grid, kGroup, 0, st>>>(a...);
        break;
    case MCKL_HIP_SYSTEMATIC:
        k_rs_counts<MCKL_HIP_SYSTEMATIC><<<grid, kGroup, 0, st>>>(a...);
        break;
    case MCKL_HIP_RESIDUAL:
        k_rs_counts<MCKL_HIP_RESIDUAL><<<grid, kGroup, 0, st>>>(a...);
        break;
    case MCKL_HIP_RESIDUAL_STRATIFIED:
        k_rs_counts<MCKL_HIP_RESIDUAL_STRATIFIED><<<grid, kGroup, 0, st>>>(
            a...);
        break;
    default:
        k_rs_counts<MCKL_HIP_RESIDUAL_SYSTEMATIC><<<grid, kGroup, 0, st>>>(
            a...);
        break;
    }
}

extern "C" {

int mckl_hip_smc_create(int model, const double *theta, size_t T,
    const double *h_y, size_t N, uint64_t seed, double threshold, int scheme,
    mckl_hip_smc **out)
{
    if (!out || !h_y)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    *out = nullptr;
    int rc = check_common(N, T, threshold, scheme);
    if (rc)
        return rc;
    ModelConst mc;
    rc = model_const(model, theta, mc);
    if (rc)
        return rc;
    mckl_hip_smc *s = new mckl_hip_smc();
    std::memset(&s->B, 0, sizeof(s->B));
    s->T = T;
    s->A.mc = mc;
    s->A.k0 = (uint32_t)seed;
    s->A.k1 = (uint32_t)(seed >> 32);
    s->A.N = (uint32_t)N;
    s->A.ntiles = (uint32_t)ceil_div(N, kTile);
    s->A.tauN = threshold * (double)N;
    s->A.scheme = scheme;
    s->A.T = (uint32_t)T;
    const size_t nt = s->A.ntiles;
    const size_t Np = nt * kTile; // padded: vector accesses stay in bounds
#define MCKL_ALLOC(p, n)                                                      \
    do {                                                                      \
        rc = dev_alloc(&(p), (n));                                            \
        if (rc) {                                                             \
            smc_free(s);                                                      \
            return rc;                                                        \
        }                                                                     \
    } while (0)
    MCKL_ALLOC(s->x[0], Np);
    MCKL_ALLOC(s->x[1], Np);
    MCKL_ALLOC(s->lw, Np);
    MCKL_ALLOC(s->idx, Np);
    MCKL_ALLOC(s->obs, T);
    MCKL_ALLOC(s->tmax, nt);
    MCKL_ALLOC(s->ts, nt);
    MCKL_ALLOC(s->ts2, nt);
    MCKL_ALLOC(s->B.tdh, nt);
    MCKL_ALLOC(s->B.tdl, nt);
    MCKL_ALLOC(s->B.te, nt);
    MCKL_ALLOC(s->B.tz, nt);
    MCKL_ALLOC(s->B.cnt, Np);
    MCKL_ALLOC(s->B.src, Np);
    MCKL_ALLOC(s->B.lbw, 2 * nt);
    if (is_residual(scheme)) {
        MCKL_ALLOC(s->B.ti, nt);
        MCKL_ALLOC(s->B.trh, nt);
        MCKL_ALLOC(s->B.trl, nt);
    }
    if (scheme == MCKL_HIP_MULTINOMIAL || scheme == MCKL_HIP_RESIDUAL) {
        MCKL_ALLOC(s->B.tmh, nt);
        MCKL_ALLOC(s->B.tml, nt);
        MCKL_ALLOC(s->B.u, Np);
        MCKL_ALLOC(s->B.uh, nt);
        MCKL_ALLOC(s->B.ul, nt);
    }
    MCKL_ALLOC(s->st, 1);
    MCKL_ALLOC(s->ess, T);
    MCKL_ALLOC(s->res, T);
#undef MCKL_ALLOC
    std::vector<float> ob(T);
    obs_const(model, theta, h_y, T, ob.data());
    hipError_t err = hipMemcpy(s->obs, ob.data(), T * sizeof(float),
        hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemset(s->st, 0, sizeof(SmcState));
    if (err == hipSuccess) // every parent a code may name stays below N
        err = hipMemset(s->B.src, 0, Np * sizeof(uint32_t));
    if (err == hipSuccess)
        err = hipEventCreateWithFlags(&s->done, hipEventDisableTiming);
    if (err == hipSuccess)
        err = hipStreamCreateWithFlags(&s->rs, hipStreamNonBlocking);
    if (err != hipSuccess) {
        rc = hip_status(err, "mckl_hip_smc_create");
        smc_free(s);
        return rc;
    }
    s->A.obs = s->obs;
    *out = s;
    return MCKL_HIP_OK;
}

// Enqueue the whole run: init, then per step the move with the tile sums,
// and the decision with the resampling counts and the rep->index transform
// (systematic, stratified: one kernel; multinomial and residual: their
// count kernels, then k_rs_index).  The resampling kernels exit at once on
// steps that do not resample.
static int smc_enqueue(mckl_hip_smc *s, hipStream_t st)
{
    const dim3 grid(s->A.ntiles);
    k_init<<<grid, kGroup, 0, st>>>(s->A, s->x[0], s->lw, s->st, s->B);
    MCKL_HIP_CHECK_LAUNCH("k_init");
    for (uint32_t t = 1; t <= s->T; ++t) {
        const float *xin = s->x[(t - 1) & 1];
        float *xout = s->x[t & 1];
        k_move<<<grid, kGroup, 0, st>>>(s->A, t, xin, xout, s->lw, s->idx,
            s->st, s->tmax, s->ts, s->ts2, s->B);
        if (s->kmom)
            k_est_moments<<<grid, kGroup, 0, st>>>(s->A, t, xout, s->lw,
                s->tmax, s->kmom, s->eph, s->epl);
        launch_rs_counts(s->A.scheme, grid, st, s->A, t, s->lw, s->tmax,
            s->ts, s->ts2, s->B, s->st, s->ess, s->res, s->idx);
        if (s->A.scheme == MCKL_HIP_MULTINOMIAL) {
            k_rs_uniforms<<<grid, kGroup, 0, st>>>(s->A, t, s->st, s->B);
            k_rs_mcounts<<<grid, kGroup, 0, st>>>(
                s->A, t, s->lw, s->tmax, s->st, s->B);
        } else if (is_residual(s->A.scheme)) {
            if (s->A.scheme == MCKL_HIP_RESIDUAL) {
                k_rs_rterms<<<grid, kGroup, 0, st>>>(s->A, t, s->st, s->B);
                k_rs_uniforms<<<grid, kGroup, 0, st>>>(s->A, t, s->st, s->B);
            }
            k_rs_rcounts<<<grid, kGroup, 0, st>>>(
                s->A, t, s->lw, s->tmax, s->st, s->B);
        }
        if (s->A.scheme == MCKL_HIP_MULTINOMIAL || is_residual(s->A.scheme))
            k_rs_index<<<grid, kGroup, 0, st>>>(s->A, t, s->st, s->B, s->idx);
    }
    if (s->kmom)
        k_est_moments_final<<<(uint32_t)s->T, kGroup, 0, st>>>(
            s->A.ntiles, s->kmom, s->eph, s->epl, s->est);
    MCKL_HIP_CHECK_LAUNCH("smc step kernels");
    return MCKL_HIP_OK;
}

int mckl_hip_smc_estimate_moments(mckl_hip_smc *s, int k)
{
    if (!s)
        return set_error(MCKL_HIP_EINVAL, "sampler is NULL");
    if (k < 0 || k > kMaxMoments)
        return set_error(MCKL_HIP_EINVAL,
            "moments: k = %d outside 0..%d", k, kMaxMoments);
    MCKL_HIP_TRY(hipEventSynchronize(s->done));
    smc_drop_graph(s);
    for (double **p : {&s->eph, &s->epl, &s->est}) {
        if (*p)
            MCKL_HIP_TRY(hipFree(*p));
        *p = nullptr;
    }
    s->kmom = 0;
    if (k == 0)
        return MCKL_HIP_OK;
    const size_t np = s->T * (size_t)(k + 1) * s->A.ntiles;
    int rc = dev_alloc(&s->eph, np);
    if (rc == MCKL_HIP_OK)
        rc = dev_alloc(&s->epl, np);
    if (rc == MCKL_HIP_OK)
        rc = dev_alloc(&s->est, s->T * (size_t)k);
    if (rc)
        return rc;
    s->kmom = k;
    return MCKL_HIP_OK;
}

int mckl_hip_smc_estimates(mckl_hip_smc *s, double *h_est)
{
    if (!s || !h_est)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (!s->kmom)
        return set_error(MCKL_HIP_EINVAL,
            "no estimator: call mckl_hip_smc_estimate_moments first");
    MCKL_HIP_TRY(hipEventSynchronize(s->done));
    MCKL_HIP_TRY(hipMemcpyWithStream(h_est, s->est,
        s->T * s->kmom * sizeof(double), hipMemcpyDeviceToHost, s->rs));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_run(mckl_hip_smc *s, void *stream)
{
    if (!s)
        return set_error(MCKL_HIP_EINVAL, "sampler is NULL");
    hipStream_t st = as_stream(stream);
    // The T-step launch sequence is captured once into a HIP graph and
    // replayed; a run contains no host synchronisation.
    if (!s->exec) {
        hipStream_t cap;
        MCKL_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t err =
            hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int rc = MCKL_HIP_OK;
        if (err == hipSuccess) {
            rc = smc_enqueue(s, cap);
            hipGraph_t g = nullptr;
            err = hipStreamEndCapture(cap, &g);
            s->graph = g;
        }
        if (err == hipSuccess && rc == MCKL_HIP_OK)
            err = hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
        if (rc)
            return rc;
        MCKL_HIP_TRY(err);
    }
    MCKL_HIP_TRY(hipGraphLaunch(s->exec, st));
    MCKL_HIP_TRY(hipEventRecord(s->done, st));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_result(mckl_hip_smc *s, double *h_loglik, double *h_ess,
    int32_t *h_resampled)
{
    if (!s)
        return set_error(MCKL_HIP_EINVAL, "sampler is NULL");
    MCKL_HIP_TRY(hipEventSynchronize(s->done));
    SmcState hs;
    MCKL_HIP_TRY(hipMemcpyWithStream(&hs, s->st, sizeof(hs),
        hipMemcpyDeviceToHost, s->rs));
    if (hs.err)
        return set_error(MCKL_HIP_EHIP,
            "smc run: a resampling look-back wait gave up; results invalid");
    if (h_loglik)
        *h_loglik = hs.logz;
    if (h_ess)
        MCKL_HIP_TRY(hipMemcpyWithStream(h_ess, s->ess, s->T * sizeof(double),
            hipMemcpyDeviceToHost, s->rs));
    if (h_resampled)
        MCKL_HIP_TRY(hipMemcpyWithStream(h_resampled, s->res,
            s->T * sizeof(int32_t), hipMemcpyDeviceToHost, s->rs));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_destroy(mckl_hip_smc *s)
{
    if (s)
        smc_free(s);
    return MCKL_HIP_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// batched sampler host side
// ---------------------------------------------------------------------------
struct mckl_hip_smc_batch {
    size_t nfilters = 0;
    size_t T = 0;
    FilterDesc *fd = nullptr;
    uint32_t *order = nullptr;
    uint32_t *queue = nullptr;
    float *xa = nullptr, *xb = nullptr, *lw = nullptr, *obs = nullptr;
    uint32_t *idx = nullptr, *cnt = nullptr, *E = nullptr, *zpos = nullptr;
    double *u = nullptr; // multinomial uniforms
    double *logz = nullptr;
    mckl_hip_tune tune{256, 0, 0, 256};
    // filters order[0, nlarge) run in the launch of tune.block-thread
    // workgroups (`persistent` of them), the rest beside it in one of
    // tune.small_block threads (`persistent_small`)
    uint32_t nlarge = 0;
    int persistent = 0;
    int persistent_small = 0;
    hipStream_t side = nullptr;              // the second launch's stream
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<uint32_t> nsorted;           // N in queue order
    double threshold = 0.0;
    int scheme = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    long long graph_tune = -1;
    // recorded on the caller's stream after each run, and the stream the
    // results are read through (see mckl_hip_smc)
    hipEvent_t done = nullptr;
    hipStream_t rs = nullptr;
    // host copies for set_params: descriptors, model, observations
    int model = 0;
    std::vector<FilterDesc> fdh;
    std::vector<double> y;
};

static void batch_free(mckl_hip_smc_batch *b)
{
    if (b->side)
        (void)hipStreamDestroy(b->side);
    if (b->fork)
        (void)hipEventDestroy(b->fork);
    if (b->join)
        (void)hipEventDestroy(b->join);
    if (b->done)
        (void)hipEventDestroy(b->done);
    if (b->rs)
        (void)hipStreamDestroy(b->rs);
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    void *ptrs[] = {b->fd, b->order, b->queue, b->xa, b->xb, b->lw, b->obs,
        b->idx, b->cnt, b->E, b->zpos, b->u, b->logz};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    delete b;
}

static void launch_k_batch(int G, uint32_t grid, hipStream_t st,
    const mckl_hip_smc_batch *b, const uint32_t *order, uint32_t nf,
    uint32_t *queue)
{
    switch (G) {
        case 1:
            k_batch<1><<<grid, kGroup, 0, st>>>(b->fd, order, nf, queue,
                b->xa, b->xb, b->lw, b->idx, b->cnt, b->E, b->zpos, b->u,
                b->logz);
            break;
        case 2:
            k_batch<2><<<grid, 2 * kGroup, 0, st>>>(b->fd, order, nf, queue,
                b->xa, b->xb, b->lw, b->idx, b->cnt, b->E, b->zpos, b->u,
                b->logz);
            break;
        default:
            k_batch<4><<<grid, 4 * kGroup, 0, st>>>(b->fd, order, nf, queue,
                b->xa, b->xb, b->lw, b->idx, b->cnt, b->E, b->zpos, b->u,
                b->logz);
            break;
    }
}

// The persistent launch(es) of a run: the large filters on `st`, the others
// on the handle's side stream, forked from and joined back into `st` (both
// captured into the same graph).  Each launch takes its filters from its
// own queue counter; every workgroup exits when its queue is empty.
static int batch_launch(mckl_hip_smc_batch *b, hipStream_t st)
{
    k_zero_u32<<<1, 64, 0, st>>>(b->queue, 2);
    const uint32_t nf = (uint32_t)b->nfilters;
    const uint32_t nl = b->nlarge;
    if (nl > 0)
        launch_k_batch(std::max(1, b->tune.block / kGroup),
            (uint32_t)b->persistent, st, b, b->order, nl, b->queue);
    if (nl < nf) {
        MCKL_HIP_TRY(hipEventRecord(b->fork, st));
        MCKL_HIP_TRY(hipStreamWaitEvent(b->side, b->fork, 0));
        launch_k_batch(std::max(1, b->tune.small_block / kGroup),
            (uint32_t)b->persistent_small, b->side, b, b->order + nl, nf - nl,
            b->queue + 1);
        MCKL_HIP_TRY(hipEventRecord(b->join, b->side));
        MCKL_HIP_TRY(hipStreamWaitEvent(st, b->join, 0));
    }
    MCKL_HIP_CHECK_LAUNCH("k_batch");
    return MCKL_HIP_OK;
}

// resident workgroups per CU of k_batch<G>, capped by waves_per_eu
static int batch_per_cu(int G, int waves_per_eu, int &per_cu)
{
    int occ = 0;
    hipError_t err;
    switch (G) {
        case 1:
            err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, k_batch<1>, kGroup, 0);
            break;
        case 2:
            err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, k_batch<2>, 2 * kGroup, 0);
            break;
        default:
            err = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, k_batch<4>, 4 * kGroup, 0);
            break;
    }
    MCKL_HIP_TRY(err);
    per_cu = std::max(1, occ);
    if (waves_per_eu > 0)
        per_cu = std::min(per_cu,
            std::max(1, waves_per_eu * 4 / (G * kGroup / 64)));
    return MCKL_HIP_OK;
}

static int set_persistent(mckl_hip_smc_batch *b)
{
    int dev = 0, ncu = 0;
    MCKL_HIP_TRY(hipGetDevice(&dev));
    MCKL_HIP_TRY(hipDeviceGetAttribute(
        &ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t nf = (uint32_t)b->nfilters;
    uint32_t nl = nf;
    if (b->tune.split_tiles > 0) {
        nl = 0;
        while (nl < nf &&
            ceil_div(b->nsorted[nl], kTile) >= (uint64_t)b->tune.split_tiles)
            ++nl;
    }
    b->nlarge = nl;
    int pl = 0, ps = 0;
    int rc = batch_per_cu(std::max(1, b->tune.block / kGroup),
        b->tune.waves_per_eu, pl);
    if (rc == MCKL_HIP_OK)
        rc = batch_per_cu(std::max(1, b->tune.small_block / kGroup),
            b->tune.waves_per_eu, ps);
    if (rc)
        return rc;
    b->persistent = (int)std::min<size_t>((size_t)ncu * pl, nl);
    b->persistent_small = (int)std::min<size_t>((size_t)ncu * ps, nf - nl);
    return MCKL_HIP_OK;
}

extern "C" {

int mckl_hip_smc_batch_create(int model, size_t nfilters, const uint32_t *h_N,
    const double *h_theta, size_t T, const double *h_y, const uint64_t *h_seed,
    double threshold, int scheme, mckl_hip_smc_batch **out)
{
    if (!out || !h_N || !h_theta || !h_y || !h_seed)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    *out = nullptr;
    if (nfilters == 0 || nfilters >= (1u << 31))
        return set_error(MCKL_HIP_EINVAL, "nfilters out of range");
    std::vector<FilterDesc> fd(nfilters);
    std::vector<float> obs(nfilters * T);
    uint64_t total = 0;
    for (size_t f = 0; f < nfilters; ++f) {
        const size_t N = h_N[f];
        int rc = check_common(N, T, threshold, scheme);
        if (rc)
            return rc;
        if (N > (size_t)kMaxTiles * kTile)
            return set_error(MCKL_HIP_EINVAL,
                "batched filters hold at most %d particles", kMaxTiles * kTile);
        ModelConst mc;
        rc = model_const(model, h_theta + 3 * f, mc);
        if (rc)
            return rc;
        SmcArgs &A = fd[f].A;
        A.mc = mc;
        A.k0 = (uint32_t)h_seed[f];
        A.k1 = (uint32_t)(h_seed[f] >> 32);
        A.N = (uint32_t)N;
        A.ntiles = (uint32_t)ceil_div(N, kTile);
        A.tauN = threshold * (double)N;
        A.scheme = scheme;
        A.T = (uint32_t)T;
        A.obs = nullptr; // set after upload
        fd[f].off = total;
        total += (uint64_t)A.ntiles * kTile;
        obs_const(model, h_theta + 3 * f, h_y + T * f, T, obs.data() + T * f);
    }
    // largest filters first (longest-processing-time order)
    std::vector<uint32_t> order(nfilters);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(),
        [&](uint32_t a, uint32_t b) { return h_N[a] > h_N[b]; });

    mckl_hip_smc_batch *b = new mckl_hip_smc_batch();
    b->nfilters = nfilters;
    b->T = T;
    b->model = model;
    b->y.assign(h_y, h_y + nfilters * T);
    int rc = MCKL_HIP_OK;
#define MCKL_ALLOC(p, n)                                                      \
    do {                                                                      \
        rc = dev_alloc(&(p), (n));                                            \
        if (rc) {                                                             \
            batch_free(b);                                                    \
            return rc;                                                        \
        }                                                                     \
    } while (0)
    MCKL_ALLOC(b->fd, nfilters);
    MCKL_ALLOC(b->order, nfilters);
    MCKL_ALLOC(b->queue, 2);
    MCKL_ALLOC(b->xa, total);
    MCKL_ALLOC(b->xb, total);
    MCKL_ALLOC(b->lw, total);
    MCKL_ALLOC(b->idx, total);
    MCKL_ALLOC(b->cnt, total);
    MCKL_ALLOC(b->E, total);
    MCKL_ALLOC(b->zpos, total);
    if (scheme == MCKL_HIP_MULTINOMIAL || scheme == MCKL_HIP_RESIDUAL)
        MCKL_ALLOC(b->u, total);
    MCKL_ALLOC(b->obs, nfilters * T);
    MCKL_ALLOC(b->logz, nfilters);
#undef MCKL_ALLOC
    for (size_t f = 0; f < nfilters; ++f)
        fd[f].A.obs = b->obs + T * f;
    hipError_t err = hipMemcpy(b->fd, fd.data(), nfilters * sizeof(FilterDesc),
        hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(b->order, order.data(), nfilters * sizeof(uint32_t),
            hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(b->obs, obs.data(), obs.size() * sizeof(float),
            hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        rc = hip_status(err, "mckl_hip_smc_batch_create");
        batch_free(b);
        return rc;
    }
    b->nsorted.resize(nfilters);
    for (size_t i = 0; i < nfilters; ++i)
        b->nsorted[i] = h_N[order[i]];
    b->threshold = threshold;
    b->scheme = scheme;
    rc = set_persistent(b);
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipEventCreateWithFlags(&b->done,
            hipEventDisableTiming), "mckl_hip_smc_batch_create");
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipEventCreateWithFlags(&b->fork,
            hipEventDisableTiming), "mckl_hip_smc_batch_create");
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipEventCreateWithFlags(&b->join,
            hipEventDisableTiming), "mckl_hip_smc_batch_create");
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipStreamCreateWithFlags(&b->side,
            hipStreamNonBlocking), "mckl_hip_smc_batch_create");
    if (rc == MCKL_HIP_OK)
        rc = hip_status(hipStreamCreateWithFlags(&b->rs,
            hipStreamNonBlocking), "mckl_hip_smc_batch_create");
    if (rc) {
        batch_free(b);
        return rc;
    }
    b->fdh = std::move(fd);
    *out = b;
    return MCKL_HIP_OK;
}

int mckl_hip_smc_batch_set_params(mckl_hip_smc_batch *b,
    const double *h_theta, const uint64_t *h_seed)
{
    if (!b || !h_theta || !h_seed)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    const size_t nf = b->nfilters, T = b->T;
    std::vector<ModelConst> mc(nf);
    for (size_t f = 0; f < nf; ++f) {
        const int rc = model_const(b->model, h_theta + 3 * f, mc[f]);
        if (rc)
            return rc;
    }
    std::vector<float> obs(nf * T);
    for (size_t f = 0; f < nf; ++f) {
        SmcArgs &A = b->fdh[f].A;
        A.mc = mc[f];
        A.k0 = (uint32_t)h_seed[f];
        A.k1 = (uint32_t)(h_seed[f] >> 32);
        obs_const(b->model, h_theta + 3 * f, b->y.data() + T * f, T,
            obs.data() + T * f);
    }
    // a run still in flight reads the descriptors: wait for it
    MCKL_HIP_TRY(hipEventSynchronize(b->done));
    MCKL_HIP_TRY(hipMemcpyWithStream(b->fd, b->fdh.data(),
        nf * sizeof(FilterDesc), hipMemcpyHostToDevice, b->rs));
    MCKL_HIP_TRY(hipMemcpyWithStream(b->obs, obs.data(),
        obs.size() * sizeof(float), hipMemcpyHostToDevice, b->rs));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_batch_set_tune(mckl_hip_smc_batch *b, const mckl_hip_tune *t)
{
    if (!b || !t)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (t->block != 256 && t->block != 512 && t->block != 1024)
        return set_error(MCKL_HIP_EINVAL, "block must be 256, 512 or 1024");
    if (t->waves_per_eu < 0 || t->waves_per_eu > 8)
        return set_error(MCKL_HIP_EINVAL, "waves_per_eu must be in 0..8");
    if (t->split_tiles < 0 || t->split_tiles > kMaxTiles)
        return set_error(MCKL_HIP_EINVAL, "split_tiles must be in 0..%d",
            kMaxTiles);
    if (t->split_tiles > 0 && t->small_block != 256 &&
        t->small_block != 512 && t->small_block != 1024)
        return set_error(MCKL_HIP_EINVAL,
            "small_block must be 256, 512 or 1024");
    MCKL_HIP_TRY(hipEventSynchronize(b->done)); // a run in flight uses it
    b->tune = *t;
    if (b->tune.split_tiles == 0)
        b->tune.small_block = 256;
    return set_persistent(b);
}

int mckl_hip_smc_batch_run(mckl_hip_smc_batch *b, int use_graph, void *stream)
{
    if (!b)
        return set_error(MCKL_HIP_EINVAL, "batch is NULL");
    hipStream_t st = as_stream(stream);
    if (!use_graph) {
        const int rc = batch_launch(b, st);
        if (rc)
            return rc;
        MCKL_HIP_TRY(hipEventRecord(b->done, st));
        return MCKL_HIP_OK;
    }
    const long long key =
        (((long long)b->tune.block * 16 + b->tune.waves_per_eu) * 128 +
            b->tune.split_tiles) * 4096 + b->tune.small_block;
    if (!b->exec || b->graph_tune != key) {
        if (b->exec) {
            (void)hipGraphExecDestroy(b->exec);
            b->exec = nullptr;
        }
        if (b->graph) {
            (void)hipGraphDestroy(b->graph);
            b->graph = nullptr;
        }
        hipStream_t cap;
        MCKL_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t err = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        int rc = MCKL_HIP_OK;
        if (err == hipSuccess) {
            rc = batch_launch(b, cap);
            err = hipStreamEndCapture(cap, &b->graph);
        }
        if (err == hipSuccess && rc == MCKL_HIP_OK)
            err = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cap);
        if (rc)
            return rc;
        MCKL_HIP_TRY(err);
        b->graph_tune = key;
    }
    MCKL_HIP_TRY(hipGraphLaunch(b->exec, st));
    MCKL_HIP_TRY(hipEventRecord(b->done, st));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_batch_result(mckl_hip_smc_batch *b, double *h_loglik)
{
    if (!b || !h_loglik)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    MCKL_HIP_TRY(hipEventSynchronize(b->done));
    MCKL_HIP_TRY(hipMemcpyWithStream(h_loglik, b->logz,
        b->nfilters * sizeof(double), hipMemcpyDeviceToHost, b->rs));
    return MCKL_HIP_OK;
}

} // extern "C"

// ---- autotuner ----
// The launch configurations tried (none changes a bit of the results: every
// filter runs through the same tile stages whatever the workgroup width,
// occupancy or launch it lands in).
static const mckl_hip_tune kTuneCandidates[] = {
    {256, 0, 0, 256}, {256, 2, 0, 256}, {512, 0, 0, 256}, {512, 2, 0, 256},
    {1024, 0, 0, 256}, {512, 0, 2, 256}, {512, 0, 4, 256}, {512, 0, 8, 256},
    {1024, 0, 4, 256}, {1024, 0, 8, 256}, {1024, 0, 16, 256},
    {1024, 0, 24, 256}, {512, 2, 8, 256}, {1024, 2, 8, 256},
};

// Workload signature for the cache: device, model, scheme, T, threshold,
// number of filters and an FNV-1a hash of their sizes in queue order.
static std::string batch_signature(const mckl_hip_smc_batch *b)
{
    int dev = 0;
    hipDeviceProp_t prop;
    std::memset(&prop, 0, sizeof(prop));
    if (hipGetDevice(&dev) != hipSuccess ||
        hipGetDeviceProperties(&prop, dev) != hipSuccess)
        std::strcpy(prop.gcnArchName, "unknown");
    uint64_t h = 1469598103934665603ull;
    for (uint32_t n : b->nsorted) {
        for (int k = 0; k < 4; ++k) {
            h ^= (n >> (8 * k)) & 0xFFu;
            h *= 1099511628211ull;
        }
    }
    char buf[512];
    std::snprintf(buf, sizeof(buf),
        "%s/%dcu/model%d/scheme%d/T%zu/thr%.17g/nf%zu/N%016llx",
        prop.gcnArchName, prop.multiProcessorCount, b->model, b->scheme, b->T,
        b->threshold, b->nfilters, (unsigned long long)h);
    return std::string(buf);
}

static bool json_int(const std::string &line, const char *name, int &v)
{
    const std::string pat = std::string("\"") + name + "\":";
    const size_t p = line.find(pat);
    if (p == std::string::npos)
        return false;
    v = std::atoi(line.c_str() + p + pat.size());
    return true;
}

// the cached configuration of `key` (the last record for it), if any
static bool cache_lookup(const char *path, const std::string &key,
    mckl_hip_tune &t, double &ms)
{
    FILE *f = std::fopen(path, "r");
    if (!f)
        return false;
    bool found = false;
    char buf[2048];
    const std::string kpat = "\"key\": \"" + key + "\"";
    while (std::fgets(buf, sizeof(buf), f)) {
        const std::string line(buf);
        if (line.find(kpat) == std::string::npos)
            continue;
        mckl_hip_tune c{256, 0, 0, 256};
        if (!json_int(line, "block", c.block) ||
            !json_int(line, "waves_per_eu", c.waves_per_eu) ||
            !json_int(line, "split_tiles", c.split_tiles) ||
            !json_int(line, "small_block", c.small_block))
            continue;
        const size_t p = line.find("\"ms\":");
        ms = p == std::string::npos ? 0.0 : std::atof(line.c_str() + p + 5);
        t = c;
        found = true;
    }
    std::fclose(f);
    return found;
}

static void cache_store(const char *path, const std::string &key,
    const mckl_hip_tune &t, double ms)
{
    FILE *f = std::fopen(path, "a");
    if (!f)
        return;
    std::fprintf(f,
        "{\"key\": \"%s\", \"block\": %d, \"waves_per_eu\": %d, "
        "\"split_tiles\": %d, \"small_block\": %d, \"ms\": %.4f}\n",
        key.c_str(), t.block, t.waves_per_eu, t.split_tiles, t.small_block,
        ms);
    std::fclose(f);
}

// one graph run timed with events on st (after a warm-up run that also
// captures the configuration's graph); the filters' log-likelihoods to h_ll
static int time_run(mckl_hip_smc_batch *b, hipStream_t st, double *h_ll,
    double &ms)
{
    int rc = mckl_hip_smc_batch_run(b, 1, st);
    if (rc)
        return rc;
    hipEvent_t e0, e1;
    MCKL_HIP_TRY(hipEventCreate(&e0));
    hipError_t err = hipEventCreate(&e1);
    if (err != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_status(err, "hipEventCreate");
    }
    err = hipEventRecord(e0, st);
    if (err == hipSuccess) {
        rc = mckl_hip_smc_batch_run(b, 1, st);
        if (rc == MCKL_HIP_OK)
            err = hipEventRecord(e1, st);
    }
    if (rc == MCKL_HIP_OK && err == hipSuccess)
        err = hipEventSynchronize(e1);
    float f = 0.0f;
    if (rc == MCKL_HIP_OK && err == hipSuccess)
        err = hipEventElapsedTime(&f, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc)
        return rc;
    MCKL_HIP_TRY(err);
    ms = f;
    return mckl_hip_smc_batch_result(b, h_ll);
}

extern "C" {

int mckl_hip_smc_batch_autotune(mckl_hip_smc_batch *b, const char *cache_path,
    mckl_hip_tune *best, double *best_ms, void *stream)
{
    if (!b)
        return set_error(MCKL_HIP_EINVAL, "batch is NULL");
    hipStream_t st = as_stream(stream);
    const std::string key = batch_signature(b);
    mckl_hip_tune t{256, 0, 0, 256};
    double ms = 0.0;
    if (cache_path && cache_lookup(cache_path, key, t, ms)) {
        const int rc = mckl_hip_smc_batch_set_tune(b, &t);
        if (rc)
            return rc;
    } else {
        std::vector<double> ref(b->nfilters), got(b->nfilters);
        bool have = false;
        for (const mckl_hip_tune &c : kTuneCandidates) {
            int rc = mckl_hip_smc_batch_set_tune(b, &c);
            double cms = 0.0;
            if (rc == MCKL_HIP_OK)
                rc = time_run(b, st, have ? got.data() : ref.data(), cms);
            if (rc)
                return rc;
            if (have && std::memcmp(got.data(), ref.data(),
                            b->nfilters * sizeof(double)) != 0)
                return set_error(MCKL_HIP_EHIP,
                    "autotune: configuration (%d, %d, %d, %d) changed the "
                    "results", c.block, c.waves_per_eu, c.split_tiles,
                    c.small_block);
            if (!have || cms < ms) {
                t = c;
                ms = cms;
            }
            have = true;
        }
        const int rc = mckl_hip_smc_batch_set_tune(b, &t);
        if (rc)
            return rc;
        if (cache_path)
            cache_store(cache_path, key, t, ms);
    }
    if (best)
        *best = t;
    if (best_ms)
        *best_ms = ms;
    return MCKL_HIP_OK;
}

int mckl_hip_smc_batch_destroy(mckl_hip_smc_batch *b)
{
    if (b)
        batch_free(b);
    return MCKL_HIP_OK;
}

} // extern "C"
