// smc_kernels.hpp -- the kernels of the fused SMC samplers, as templates on
// the model (see the model concept in smc_tiles.hpp):
//
//  * plain sampler: one workgroup per 2048-particle tile, two launches per
//    step (k_move, k_rs_counts<scheme>) with no host synchronisation (the
//    ESS test is taken on the device and the resampling kernels exit when
//    the step does not resample), so the whole T loop can be captured into
//    one HIP graph;
//  * batched sampler: thousands of independent filters of uneven sizes in
//    one persistent launch, one workgroup per filter taken from a device
//    queue ordered by decreasing particle count;
// both built from the per-tile stages of smc_tiles.hpp, so a filter gives
// the same bits in either sampler.  The library instantiates them on the
// built-in models (smc.hip); a program compiled with hipcc instantiates them
// on its own model through mckl_amd/include/mckl_hip_fused.hpp.
#pragma once

#include "common.hpp"
#include "smc_tiles.hpp"

namespace mckl_hip {
// Internal linkage: every translation unit that includes this header (the
// library's smc.hip, a program's model) has kernels and types of its own, so
// two modules never share a kernel symbol.
namespace {

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
    double *tpre;      // systematic/stratified: each thread's exclusive dd
                       // prefix of a within its tile (hi, lo), from k_move
};

__device__ __forceinline__ void store_dd_nt(double *p, size_t i, dd v)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 w = {v.hi, v.lo};
    __builtin_nontemporal_store(w, reinterpret_cast<d2 *>(p) + i);
}

__device__ __forceinline__ dd load_dd(const double *p, size_t i)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 w = reinterpret_cast<const d2 *>(p)[i];
    return dd{w.x, w.y};
}

template <class M>
__global__ __launch_bounds__(kGroup) void k_init(
    SmcArgs A, M model, float *x, float *lw, SmcState *st, RsBuf B)
{
    tile_init(A, model, blockIdx.x, x, lw, threadIdx.x);
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
// is applied by the next kernel); with B.tpre (systematic, stratified) each
// thread also stores its exclusive double-double prefix of a within the
// tile, which k_rs_counts would otherwise recompute
template <class M>
__global__ __launch_bounds__(kGroup) void k_move(SmcArgs A, M model,
    uint32_t t, const float *xin, float *xout, float *lw, const uint32_t *idx,
    const SmcState *st, float *tmax, double *ts, double *ts2, RsBuf B)
{
    __shared__ __align__(16) unsigned char lds[kScratch];
    const GroupLds L = group_lds(lds);
    const int p = (t - 1) & 1;
    float l[kItems];
    const float mt = tile_move<true>(A, model, t, blockIdx.x, xin, xout, lw,
        idx, B.src, st->resampled[p], st->m[p], threadIdx.x, L.f, l);
    double S, S2;
    dd et;
    const dd lp =
        tile_sums_l(A, blockIdx.x, l, mt, threadIdx.x, L.d, S, S2, et);
    if (B.tpre)
        store_dd_nt(B.tpre, (size_t)blockIdx.x * kGroup + threadIdx.x, lp);
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
#pragma unroll 4
    for (uint32_t i = gt; i < nt; i += kGroup)
        m = fmaxf(m, tmax[i]);
    return group_max(m, gt, ldsf);
}

// Every workgroup reduces the tile sums and takes the step's decision;
// workgroup 0 records it.  When the step resamples, each workgroup derives
// its tile's prefix offsets and counts the uniforms below its cumulative
// weights.  All global loads the workgroup needs are issued up front (its
// tile's log weights and thread prefixes, its chunk of every per-tile
// array), so the kernel waits on memory once; the reductions are those of
// tiles_max, tiles_sum2 and tiles_offset_dd, operation for operation.
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
    constexpr bool kSorted =
        S == MCKL_HIP_SYSTEMATIC || S == MCKL_HIP_STRATIFIED;
    float l[kItems];
    load8(lw, tile_first(k, gt), A.N, l, 0.0f);
    dd ttoff{0.0, 0.0}; // the thread's prefix of a within the tile
    if constexpr (kSorted)
        ttoff = load_dd(B.tpre, (size_t)k * kGroup + gt);
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
        // loads of four iterations in flight together (unrolled)
#pragma unroll 4
        for (int i = gt; i < nt; i += kGroup)
            mx = fmaxf(mx, tmax[i]);
    }
    // the step's uniform sequence (its key work overlaps the loads)
    UniformSeq U{};
    if constexpr (kSorted)
        U = make_uniform_seq(A, t);
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
#pragma unroll 4
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
    float ta[kItems]; // a_i = exp_f(lw_i - m_k)
    if constexpr (kSorted)
        tile_a(A, k, l, tk, gt, ta);
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
#pragma unroll 4
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
    uint32_t es, zs, cs[kItems], prev;
    TileScan sc;
    tile_counts_c(A, k, ta, ttoff, rk, invS, toff, next_hi, U, gt, L.u, cs,
        prev, es, zs, &sc);
    uint64_t *w = B.lbw + (size_t)c * nt;
    if (gt == 0)
        lb_publish(w, k, es, zs);
    TileIndex ti;
    tile_index_local(A, k, cs, prev, gt, L.u, Es, Srcl, ti, &sc);
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
#pragma unroll 4
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
#pragma unroll 4
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

template <class M>
struct FilterDescT {
    SmcArgs A;
    M model;
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

template <int G, class M>
__global__ __launch_bounds__(G *kGroup, G == 1 ? 2 : 1) void k_batch(
    const FilterDescT<M> *fd,
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
        const M model = fd[f].model;
        const uint64_t off = fd[f].off;
        // states by component: model_dim blocks of the filter's padded
        // size (its share of xa/xb is model_dim times its offset)
        constexpr uint64_t D = model_dim<M>::value;
        float *x0 = xa + D * off, *x1 = xb + D * off, *lw = lwb + off;
        uint32_t *idx = idxb + off, *cnt = cntb + off, *Eincl = Eb + off;
        uint32_t *zpos = zb + off;
        const int nt = (int)A.ntiles;
        const int rounds = (nt + G - 1) / G;

        for (int k = grp; k < nt; k += G)
            tile_init(A, model, (uint32_t)k, x0, lw, gt);
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
                const float mt = tile_move<false>(A, model, t, kk, xin, xout,
                    lw, idx, nullptr, resampled, mprev, gt, L.f, l);
                // the double-double totals only when the step resamples
                double s, s2;
                tile_sums2(A, kk, l, mt, gt, L.d, s, s2);
                if (gt == 0 && k < nt) {
                    S.tmax[k] = mt;
                    S.ts[k] = s;
                    S.ts2[k] = s2;
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
                // the tiles' double-double totals of a (tile_sums_l's etot,
                // from the stored log weights), scaled by r_k
                for (int r = 0; r < rounds; ++r) {
                    const int k = r * G + grp;
                    const uint32_t kk = (uint32_t)(k < nt ? k : nt);
                    float l[kItems];
                    load8(lw, tile_first(kk, gt), A.N, l, 0.0f);
                    const dd et = tile_dd_total(
                        A, kk, l, k < nt ? S.tmax[k] : 0.0f, gt, L.d);
                    if (gt == 0 && k < nt) {
                        const dd u = dd_mul_d(et, S.r[k]);
                        S.tdh[k] = u.hi;
                        S.tdl[k] = u.lo;
                    }
                    __syncthreads();
                }
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

} // namespace
} // namespace mckl_hip
