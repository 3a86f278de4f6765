// smc_math.hpp -- device building blocks of the SMC samplers.
//
//  * float log/exp/sin-cos(2 pi u) with explicit fused multiply-adds in a
//    fixed order: the same definitions as the CPU oracle, so both compute
//    identical bits (no dependence on a vendor math library);
//  * double-double (TwoSum) accumulation for the FP64 prefix sums;
//  * the canonical per-tile reductions and scans.  A tile is 2048 particles
//    handled by one group of 256 threads (4 wavefronts), 8 contiguous
//    particles per thread.  Every sampler (plain, batched) reduces and scans
//    through these functions and combines tiles with tiles_sum2 /
//    tiles_scan_*, whose order depends only on the tile count, so results do
//    not depend on how many tiles a launch processes concurrently.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engines.hpp"

namespace mckl_hip {

constexpr int kGroup = 256;          // threads per tile group
constexpr int kItems = 8;            // particles per thread
constexpr int kTile = kGroup * kItems; // particles per tile
constexpr int kWaves = kGroup / 64;

__device__ __forceinline__ float f_as(uint32_t b) { return __uint_as_float(b); }
__device__ __forceinline__ uint32_t u_as(float f) { return __float_as_uint(f); }

// log(x), x positive normal.  x = 2^e m, m in [sqrt(1/2), sqrt(2)),
// log(1+f) = f + f^2 P(f).
__device__ __forceinline__ float smc_log_f(float x)
{
    const uint32_t b = u_as(x);
    int e = (int)(b >> 23) - 127;
    uint32_t mb = (b & 0x7FFFFFu) | 0x3F800000u;
    if (mb > 0x3FB504F3u) {
        mb -= 0x00800000u;
        e += 1;
    }
    const float f = __fsub_rn(f_as(mb), 1.0f);
    float p = -0x1.4237fcp-4f;
    p = __fmaf_rn(p, f, 0x1.0696e4p-3f);
    p = __fmaf_rn(p, f, -0x1.0c524cp-3f);
    p = __fmaf_rn(p, f, 0x1.22973ap-3f);
    p = __fmaf_rn(p, f, -0x1.548882p-3f);
    p = __fmaf_rn(p, f, 0x1.99a3ecp-3f);
    p = __fmaf_rn(p, f, -0x1.000206p-2f);
    p = __fmaf_rn(p, f, 0x1.55554ep-2f);
    p = __fmaf_rn(p, f, -0x1.fffffep-2f);
    const float f2 = __fmul_rn(f, f);
    const float r = __fmaf_rn(f2, p, f);
    const float ef = (float)e;
    return __fmaf_rn(ef, 0x1.62e430p-1f, __fmaf_rn(ef, -0x1.05c610p-29f, r));
}

// exp(x); 0 below -87 (deterministic underflow), +inf above 88.  Both
// ends are selects over the evaluated polynomial (no divergent branch).
__device__ __forceinline__ float smc_exp_f(float x)
{
    const float k = rintf(__fmul_rn(x, 0x1.715476p+0f));
    float r = __fmaf_rn(-k, 0x1.62e430p-1f, x);
    r = __fmaf_rn(-k, -0x1.05c610p-29f, r);
    float q = 0x1.a1516cp-13f;
    q = __fmaf_rn(q, r, 0x1.6d431ap-10f);
    q = __fmaf_rn(q, r, 0x1.1110c6p-7f);
    q = __fmaf_rn(q, r, 0x1.5554eap-5f);
    q = __fmaf_rn(q, r, 0x1.555556p-3f);
    q = __fmaf_rn(q, r, 0x1.000000p-1f);
    const float p = __fadd_rn(1.0f, __fmaf_rn(__fmul_rn(r, r), q, r));
    const float v = __fmul_rn(p, f_as((uint32_t)((int)k + 127) << 23));
    return x < -87.0f ? 0.0f : (x > 88.0f ? __builtin_huge_valf() : v);
}

// sin(2 pi u), cos(2 pi u), u in [0, 1]
__device__ __forceinline__ void smc_sincos2pi_f(float u, float &s, float &c)
{
    const float q = rintf(__fmul_rn(u, 4.0f));
    const float r = __fsub_rn(u, __fmul_rn(q, 0.25f));
    const float ph = __fmul_rn(r, 0x1.921fb6p+2f);
    const float p2 = __fmul_rn(ph, ph);
    float sp = 0x1.6dbd24p-19f;
    sp = __fmaf_rn(sp, p2, -0x1.a013a4p-13f);
    sp = __fmaf_rn(sp, p2, 0x1.11110ep-7f);
    sp = __fmaf_rn(sp, p2, -0x1.555556p-3f);
    const float sn = __fmaf_rn(__fmul_rn(ph, p2), sp, ph);
    float cp = -0x1.2469bap-22f;
    cp = __fmaf_rn(cp, p2, 0x1.a0125ap-16f);
    cp = __fmaf_rn(cp, p2, -0x1.6c16bap-10f);
    cp = __fmaf_rn(cp, p2, 0x1.555556p-5f);
    cp = __fmaf_rn(cp, p2, -0x1.000000p-1f);
    const float cs = __fmaf_rn(p2, cp, 1.0f);
    // quadrant q mod 4: (s, c) = (sn, cs), (cs, -sn), (-sn, -cs), (-cs, sn),
    // by selects and sign flips
    const uint32_t q3 = (uint32_t)(int)q & 3u;
    const bool swap = (q3 & 1u) != 0;
    const float bs = swap ? cs : sn;
    const float bc = swap ? sn : cs;
    s = f_as(u_as(bs) ^ ((q3 & 2u) << 30));
    c = f_as(u_as(bc) ^ (((q3 + 1u) & 2u) << 30));
}

// Normal pair of particles 2k, 2k+1 at step t: Philox4x32 block at counter
// {k, t} under key {seed lo, seed hi}; even = r cos, odd = r sin.
__device__ __forceinline__ void smc_normal_pair(
    uint32_t k0, uint32_t k1, uint64_t k, uint64_t t, float &e0, float &e1)
{
    uint32_t s0 = (uint32_t)k, s1 = (uint32_t)(k >> 32);
    uint32_t s2 = (uint32_t)t, s3 = (uint32_t)(t >> 32);
    philox4x32_10(k0, k1, s0, s1, s2, s3);
    const float u1 = u01_f32(2, s0);
    const float u2 = u01_f32(2, s1);
    const float rad = sqrt_rn(__fmul_rn(-2.0f, smc_log_f(u1)));
    float sn, cs;
    smc_sincos2pi_f(u2, sn, cs);
    e0 = __fmul_rn(rad, cs);
    e1 = __fmul_rn(rad, sn);
}

// Canonical uniform j of step t (resampling): counter {2^63 + j/2, t}
__device__ __forceinline__ double smc_uniform(
    uint32_t k0, uint32_t k1, uint64_t t, uint64_t j)
{
    const uint64_t c = 0x8000000000000000ull + (j >> 1);
    uint32_t s0 = (uint32_t)c, s1 = (uint32_t)(c >> 32);
    uint32_t s2 = (uint32_t)t, s3 = (uint32_t)(t >> 32);
    philox4x32_10(k0, k1, s0, s1, s2, s3);
    return (j & 1) ? u01_canonical_f64(s2, s3) : u01_canonical_f64(s0, s1);
}

// ---- model arithmetic (float, one rounding per operation) ----
struct ModelConst {
    int model;      // 1 LG, 2 SV
    float a, s;     // transition
    float sd0;      // initial scale
    float inv_sy;   // LG observation precision scale
    float cst;      // constant of log g
};

__device__ __forceinline__ float smc_move(const ModelConst &m, float x, float e)
{
    return __fadd_rn(__fmul_rn(m.a, x), __fmul_rn(m.s, e));
}

// obs = (float)y for LG, (float)(y^2/beta^2) for SV
__device__ __forceinline__ float smc_logg(const ModelConst &m, float obs, float x)
{
    if (m.model == 1) {
        const float d = __fsub_rn(obs, x);
        const float z = __fmul_rn(d, m.inv_sy);
        return __fadd_rn(__fmul_rn(-0.5f, __fmul_rn(z, z)), m.cst);
    }
    const float q = __fmul_rn(obs, smc_exp_f(-x));
    return __fadd_rn(__fsub_rn(__fmul_rn(-0.5f, q), __fmul_rn(0.5f, x)), m.cst);
}

// ---- double-double ----
struct dd {
    double hi, lo;
};

__device__ __forceinline__ dd dd_add(dd a, dd b)
{
    const double s = __dadd_rn(a.hi, b.hi);
    const double bb = __dsub_rn(s, a.hi);
    const double err = __dadd_rn(__dsub_rn(a.hi, __dsub_rn(s, bb)),
        __dsub_rn(b.hi, bb));
    const double lo = __dadd_rn(err, __dadd_rn(a.lo, b.lo));
    const double h = __dadd_rn(s, lo);
    return dd{h, __dsub_rn(lo, __dsub_rn(h, s))};
}

// exact product a*b = hi + lo (one fused multiply-add for the error term)
__device__ __forceinline__ dd two_prod(double a, double b)
{
    const double p = __dmul_rn(a, b);
    return dd{p, __fma_rn(a, b, -p)};
}

// a * b for dd a and double b (error about 2^-104 relative)
__device__ __forceinline__ dd dd_mul_d(dd a, double b)
{
    const dd p = two_prod(a.hi, b);
    const double lo = __fma_rn(a.lo, b, p.lo);
    const double h = __dadd_rn(p.hi, lo);
    return dd{h, __dsub_rn(lo, __dsub_rn(h, p.hi))};
}

// a / b for dd a and double b > 0, to about 2^-104 relative before the
// final rounding of hi + lo
__device__ __forceinline__ double dd_div_d(dd a, double b)
{
    const double q = __ddiv_rn(a.hi, b);
    const dd qb = two_prod(q, b);
    const double r = __dadd_rn(__dsub_rn(__dsub_rn(a.hi, qb.hi), qb.lo), a.lo);
    return __dadd_rn(q, __ddiv_rn(r, b));
}

// ---- wave (64-lane) reductions and scans on DPP lane moves ----
// A DPP move reads a register of another lane of the same row (16 lanes),
// or broadcasts lane 15 / 31 to the next rows (row_bcast, GFX9), inside a
// VALU instruction, with no LDS round trip.  Lanes without a source lane, or
// in rows outside the row mask, take the identity `idn`.  Orders are fixed:
//   reduce: quad butterflies (xor 1, xor 2), half-row and row mirrors (all
//           16 lanes of a row hold the row total), rows 0+1 and 2+3 by
//           row_bcast:15, then the four rows by row_bcast:31; lane 63 holds
//           the total, which readlane makes uniform;
//   scan:   row shifts by 1, 2, 4, 8 (inclusive within rows), then
//           row_bcast:15 and row_bcast:31 carry the row totals forward;
//           the exclusive form shifts the inclusive one by a lane.
enum : int {
    kDppXor1 = 0xB1,     // quad_perm [1, 0, 3, 2]
    kDppXor2 = 0x4E,     // quad_perm [2, 3, 0, 1]
    kDppShr1 = 0x111,    // row_shr:1
    kDppShr2 = 0x112,    // row_shr:2
    kDppShr4 = 0x114,    // row_shr:4
    kDppShr8 = 0x118,    // row_shr:8
    kDppWaveShr1 = 0x138, // wave_shr:1
    kDppMirror = 0x140,  // row_mirror
    kDppHalfMirror = 0x141, // row_half_mirror
    kDppBcast15 = 0x142, // row_bcast:15
    kDppBcast31 = 0x143, // row_bcast:31
};

template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ uint32_t dpp_u(uint32_t idn, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(
        (int)idn, (int)v, Ctrl, RowMask, 0xF, false);
}

template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ float dpp_f(float idn, float v)
{
    return __uint_as_float(dpp_u<Ctrl, RowMask>(
        __float_as_uint(idn), __float_as_uint(v)));
}

template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ double dpp_d(double idn, double v)
{
    const uint64_t a = (uint64_t)__double_as_longlong(idn);
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = dpp_u<Ctrl, RowMask>((uint32_t)a, (uint32_t)b);
    const uint32_t hi =
        dpp_u<Ctrl, RowMask>((uint32_t)(a >> 32), (uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ uint32_t lane63_u(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ float lane63_f(float v)
{
    return __uint_as_float(lane63_u(__float_as_uint(v)));
}

__device__ __forceinline__ double lane63_d(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = lane63_u((uint32_t)b);
    const uint32_t hi = lane63_u((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// wave reductions (every lane gets the result)
__device__ __forceinline__ float wave_max(float v)
{
    const float n = -__builtin_huge_valf();
    v = fmaxf(v, dpp_f<kDppXor1>(n, v));
    v = fmaxf(v, dpp_f<kDppXor2>(n, v));
    v = fmaxf(v, dpp_f<kDppHalfMirror>(n, v));
    v = fmaxf(v, dpp_f<kDppMirror>(n, v));
    v = fmaxf(v, dpp_f<kDppBcast15, 0xA>(n, v));
    v = fmaxf(v, dpp_f<kDppBcast31, 0xC>(n, v));
    return lane63_f(v);
}

__device__ __forceinline__ double dmax(double a, double b)
{
    return a < b ? b : a;
}

__device__ __forceinline__ double wave_max_d(double v)
{
    const double n = -__builtin_huge_val();
    v = dmax(v, dpp_d<kDppXor1>(n, v));
    v = dmax(v, dpp_d<kDppXor2>(n, v));
    v = dmax(v, dpp_d<kDppHalfMirror>(n, v));
    v = dmax(v, dpp_d<kDppMirror>(n, v));
    v = dmax(v, dpp_d<kDppBcast15, 0xA>(n, v));
    v = dmax(v, dpp_d<kDppBcast31, 0xC>(n, v));
    return lane63_d(v);
}

__device__ __forceinline__ double wave_sum_d(double v)
{
    v = __dadd_rn(v, dpp_d<kDppXor1>(0.0, v));
    v = __dadd_rn(v, dpp_d<kDppXor2>(0.0, v));
    v = __dadd_rn(v, dpp_d<kDppHalfMirror>(0.0, v));
    v = __dadd_rn(v, dpp_d<kDppMirror>(0.0, v));
    v = __dadd_rn(v, dpp_d<kDppBcast15, 0xA>(0.0, v));
    v = __dadd_rn(v, dpp_d<kDppBcast31, 0xC>(0.0, v));
    return lane63_d(v);
}

__device__ __forceinline__ uint32_t wave_sum_u(uint32_t v)
{
    v += dpp_u<kDppXor1>(0u, v);
    v += dpp_u<kDppXor2>(0u, v);
    v += dpp_u<kDppHalfMirror>(0u, v);
    v += dpp_u<kDppMirror>(0u, v);
    v += dpp_u<kDppBcast15, 0xA>(0u, v);
    v += dpp_u<kDppBcast31, 0xC>(0u, v);
    return lane63_u(v);
}

__device__ __forceinline__ uint32_t wave_min_u(uint32_t v)
{
    const uint32_t n = 0xFFFFFFFFu;
    v = min(v, dpp_u<kDppXor1>(n, v));
    v = min(v, dpp_u<kDppXor2>(n, v));
    v = min(v, dpp_u<kDppHalfMirror>(n, v));
    v = min(v, dpp_u<kDppMirror>(n, v));
    v = min(v, dpp_u<kDppBcast15, 0xA>(n, v));
    v = min(v, dpp_u<kDppBcast31, 0xC>(n, v));
    return lane63_u(v);
}

template <int Ctrl, int RowMask = 0xF>
__device__ __forceinline__ dd dpp_dd(dd v)
{
    return dd{dpp_d<Ctrl, RowMask>(0.0, v.hi), dpp_d<Ctrl, RowMask>(0.0, v.lo)};
}

// wave-wide dd sum (every lane gets the total)
__device__ __forceinline__ dd wave_sum_dd(dd v)
{
    v = dd_add(v, dpp_dd<kDppXor1>(v));
    v = dd_add(v, dpp_dd<kDppXor2>(v));
    v = dd_add(v, dpp_dd<kDppHalfMirror>(v));
    v = dd_add(v, dpp_dd<kDppMirror>(v));
    v = dd_add(v, dpp_dd<kDppBcast15, 0xA>(v));
    v = dd_add(v, dpp_dd<kDppBcast31, 0xC>(v));
    return dd{lane63_d(v.hi), lane63_d(v.lo)};
}

// inclusive wave scans (lane l gets the combination of lanes 0..l)
__device__ __forceinline__ uint32_t wave_incl_sum_u(uint32_t v)
{
    v += dpp_u<kDppShr1>(0u, v);
    v += dpp_u<kDppShr2>(0u, v);
    v += dpp_u<kDppShr4>(0u, v);
    v += dpp_u<kDppShr8>(0u, v);
    v += dpp_u<kDppBcast15, 0xA>(0u, v);
    v += dpp_u<kDppBcast31, 0xC>(0u, v);
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_max_u(uint32_t v)
{
    v = max(v, dpp_u<kDppShr1>(0u, v));
    v = max(v, dpp_u<kDppShr2>(0u, v));
    v = max(v, dpp_u<kDppShr4>(0u, v));
    v = max(v, dpp_u<kDppShr8>(0u, v));
    v = max(v, dpp_u<kDppBcast15, 0xA>(0u, v));
    v = max(v, dpp_u<kDppBcast31, 0xC>(0u, v));
    return v;
}

__device__ __forceinline__ dd wave_incl_sum_dd(dd v)
{
    v = dd_add(dpp_dd<kDppShr1>(v), v);
    v = dd_add(dpp_dd<kDppShr2>(v), v);
    v = dd_add(dpp_dd<kDppShr4>(v), v);
    v = dd_add(dpp_dd<kDppShr8>(v), v);
    v = dd_add(dpp_dd<kDppBcast15, 0xA>(v), v);
    v = dd_add(dpp_dd<kDppBcast31, 0xC>(v), v);
    return v;
}

// ---- group (256-thread) reductions; `lds` is this group's scratch ----
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// group max; result valid in all threads of the group (gt = thread index
// within the group, lds >= 4 floats)
__device__ __forceinline__ float group_max(float v, int gt, float *lds)
{
    v = wave_max(v);
    if ((gt & 63) == 0)
        lds[gt >> 6] = v;
    __syncthreads();
    const float r = fmaxf(fmaxf(lds[0], lds[1]), fmaxf(lds[2], lds[3]));
    __syncthreads();
    return r;
}

// group sum of two doubles in the canonical order: sequential per thread
// (done by the caller), the wave reduction above, (w0 + w1) + (w2 + w3).
__device__ __forceinline__ void group_sum2(
    double &a, double &b, int gt, double *lds)
{
    a = wave_sum_d(a);
    b = wave_sum_d(b);
    if ((gt & 63) == 0) {
        lds[gt >> 6] = a;
        lds[4 + (gt >> 6)] = b;
    }
    __syncthreads();
    a = __dadd_rn(__dadd_rn(lds[0], lds[1]), __dadd_rn(lds[2], lds[3]));
    b = __dadd_rn(__dadd_rn(lds[4], lds[5]), __dadd_rn(lds[6], lds[7]));
    __syncthreads();
}

// exclusive scan of a u32 per thread over the group; returns the exclusive
// prefix, `total` receives the group sum.  lds >= 4 u32.
__device__ __forceinline__ uint32_t group_excl_scan_u(
    uint32_t v, int gt, uint32_t *lds, uint32_t &total)
{
    const int lane = gt & 63;
    const uint32_t incl = wave_incl_sum_u(v);
    if (lane == 63)
        lds[gt >> 6] = incl;
    __syncthreads();
    uint32_t off = 0;
    const int w = gt >> 6;
    for (int i = 0; i < w; ++i)
        off += lds[i];
    total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return off + incl - v;
}

// group min of a u32; result in every thread.  lds >= 4 u32.
__device__ __forceinline__ uint32_t group_min_u(uint32_t v, int gt,
    uint32_t *lds)
{
    v = wave_min_u(v);
    if ((gt & 63) == 0)
        lds[gt >> 6] = v;
    __syncthreads();
    uint32_t r = lds[0];
    for (int w = 1; w < kWaves; ++w)
        r = lds[w] < r ? lds[w] : r;
    __syncthreads();
    return r;
}

// exclusive max-scan of a u32 per thread over the group (init for thread 0).
// lds >= 4 u32.
__device__ __forceinline__ uint32_t group_excl_max_u(
    uint32_t v, uint32_t init, int gt, uint32_t *lds)
{
    const int lane = gt & 63;
    const uint32_t incl = wave_incl_max_u(v);
    const uint32_t ex = dpp_u<kDppWaveShr1>(0u, incl); // lane 0: 0
    if (lane == 63)
        lds[gt >> 6] = incl;
    __syncthreads();
    uint32_t base = init;
    for (int w = 0; w < (gt >> 6); ++w)
        base = base > lds[w] ? base : lds[w];
    __syncthreads();
    return base > ex ? base : ex;
}

// exclusive dd scan per thread value over the group (the wave scan above,
// waves in order).  lds >= 8 doubles.
__device__ __forceinline__ dd group_excl_scan_dd(
    dd v, int gt, double *lds, dd &total)
{
    const int lane = gt & 63;
    const dd incl = wave_incl_sum_dd(v);
    // exclusive within the wave: the inclusive value of lane-1
    dd ex{dpp_d<kDppWaveShr1>(0.0, incl.hi), dpp_d<kDppWaveShr1>(0.0, incl.lo)};
    if (lane == 63) {
        lds[2 * (gt >> 6)] = incl.hi;
        lds[2 * (gt >> 6) + 1] = incl.lo;
    }
    __syncthreads();
    const int w = gt >> 6;
    dd off{0.0, 0.0};
    for (int i = 0; i < w; ++i)
        off = dd_add(off, dd{lds[2 * i], lds[2 * i + 1]});
    dd tot{0.0, 0.0};
    for (int i = 0; i < kWaves; ++i)
        tot = dd_add(tot, dd{lds[2 * i], lds[2 * i + 1]});
    total = tot;
    __syncthreads();
    return w == 0 ? ex : (lane == 0 ? off : dd_add(off, ex));
}

// ---- cross-tile combination, canonical for a given tile count ----
// Tiles are dealt to the 256 threads of one group in contiguous chunks of
// ceil(n/256); each thread combines its chunk sequentially, then the group
// combines thread results with the group functions above.  Plain (many
// workgroups) and batched (one workgroup per filter) samplers both call these
// with the same n, so they produce the same bits.

// sums of (s, s2) over n tiles; result in every thread of the group
__device__ __forceinline__ void tiles_sum2(const double *s, const double *s2,
    int n, int gt, double *lds, double &rs, double &rs2)
{
    const int chunk = (n + kGroup - 1) / kGroup;
    const int b = gt * chunk;
    const int e = b + chunk < n ? b + chunk : n;
    double a = 0.0, c = 0.0;
#pragma unroll 4
    for (int i = b; i < e; ++i) {
        a = __dadd_rn(a, s[i]);
        c = __dadd_rn(c, s2[i]);
    }
    group_sum2(a, c, gt, lds);
    rs = a;
    rs2 = c;
}

// dd total of n tile totals, the same operations as the total of
// tiles_scan_dd (chunk sums, then the group scan's total)
__device__ __forceinline__ dd tiles_total_dd(
    const double *hi, const double *lo, int n, int gt, double *lds)
{
    const int chunk = (n + kGroup - 1) / kGroup;
    const int b = gt * chunk;
    const int e = b + chunk < n ? b + chunk : n;
    dd loc{0.0, 0.0};
#pragma unroll 4
    for (int i = b; i < e; ++i)
        loc = dd_add(loc, dd{hi[i], lo[i]});
    dd total;
    (void)group_excl_scan_dd(loc, gt, lds, total);
    return total;
}

// exclusive dd scan over n tile totals (hi/lo arrays); writes off[0..n],
// off[n] = total.  Arrays may live in LDS or global memory.
__device__ __forceinline__ void tiles_scan_dd(bool write, const double *hi,
    const double *lo, int n, double *ohi, double *olo, int gt, double *lds)
{
    const int chunk = (n + kGroup - 1) / kGroup;
    const int b = gt * chunk;
    const int e = b + chunk < n ? b + chunk : n;
    dd loc{0.0, 0.0};
#pragma unroll 4
    for (int i = b; i < e; ++i)
        loc = dd_add(loc, dd{hi[i], lo[i]});
    dd total;
    dd off = group_excl_scan_dd(loc, gt, lds, total);
    if (write) {
#pragma unroll 4
        for (int i = b; i < e; ++i) {
            ohi[i] = off.hi;
            olo[i] = off.lo;
            off = dd_add(off, dd{hi[i], lo[i]});
        }
        if (gt == 0) {
            ohi[n] = total.hi;
            olo[n] = total.lo;
        }
    }
    __syncthreads();
}

// The entries j0 and j1 (0 <= j <= n) of the offsets tiles_scan_dd writes,
// computed with the same operations in the same order (the thread owning
// tile j adds its chunk's tiles before j to its exclusive prefix; j = n is
// the total), without writing the array.  Lets every workgroup of a launch
// derive its own tile offsets.  lds >= 8 doubles + 4 doubles of exchange.
__device__ __forceinline__ void tiles_offset_dd(const double *hi,
    const double *lo, int n, int j0, int j1, int gt, double *lds, dd &o0,
    dd &o1)
{
    const int chunk = (n + kGroup - 1) / kGroup;
    const int b = gt * chunk;
    const int e = b + chunk < n ? b + chunk : n;
    dd loc{0.0, 0.0};
#pragma unroll 4
    for (int i = b; i < e; ++i)
        loc = dd_add(loc, dd{hi[i], lo[i]});
    dd total;
    const dd off = group_excl_scan_dd(loc, gt, lds, total);
    double *x = lds + 8;
    if (b < n || gt == 0) {
        dd run = off;
#pragma unroll 4
        for (int i = b; i < e; ++i) {
            if (i == j0) {
                x[0] = run.hi;
                x[1] = run.lo;
            }
            if (i == j1) {
                x[2] = run.hi;
                x[3] = run.lo;
            }
            run = dd_add(run, dd{hi[i], lo[i]});
        }
    }
    if (gt == 0) {
        if (j0 == n) {
            x[0] = total.hi;
            x[1] = total.lo;
        }
        if (j1 == n) {
            x[2] = total.hi;
            x[3] = total.lo;
        }
    }
    __syncthreads();
    o0 = dd{x[0], x[1]};
    o1 = dd{x[2], x[3]};
    __syncthreads();
}

// exclusive u32 scans of two arrays over n tiles; off arrays get n+1 entries
__device__ __forceinline__ void tiles_scan_u2(bool write, const uint32_t *a,
    const uint32_t *b2, int n, uint32_t *oa, uint32_t *ob, int gt,
    uint32_t *lds)
{
    const int chunk = (n + kGroup - 1) / kGroup;
    const int b = gt * chunk;
    const int e = b + chunk < n ? b + chunk : n;
    uint32_t la = 0, lb = 0;
#pragma unroll 4
    for (int i = b; i < e; ++i) {
        la += a[i];
        lb += b2[i];
    }
    uint32_t ta, tb;
    uint32_t xa = group_excl_scan_u(la, gt, lds, ta);
    uint32_t xb = group_excl_scan_u(lb, gt, lds + 4, tb);
    if (write) {
#pragma unroll 4
        for (int i = b; i < e; ++i) {
            oa[i] = xa;
            ob[i] = xb;
            xa += a[i];
            xb += b2[i];
        }
        if (gt == 0) {
            oa[n] = ta;
            ob[n] = tb;
        }
    }
    __syncthreads();
}

} // namespace mckl_hip
