// engines.hpp -- device-side counter-based generators for gfx950.
//
// One counter block per lane.  Keys arrive as kernel arguments, which the
// compiler keeps in SGPRs, so per-round keys (key + W*r) are scalar adds.
// Reference paths are relative to the zhouyan/MCKL source tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mckl_hip {

enum { PHILOX = 1, THREEFRY = 2, ARS = 3 };

// Per-stream position handed to the fill kernels: the counter of the first
// block touched by the call and the key.  128-bit (Philox, ARS) and 256-bit
// (Threefry) counters are little-endian arrays of u64, as in
// random/increment.hpp:54-104.
struct StreamArg {
    uint64_t ctr[4];
    uint64_t key[4];
};

// counter + b with ripple carry (random/internal/increment_generic.hpp)
__device__ __forceinline__ void ctr_add2(
    const uint64_t *c, uint64_t b, uint64_t &lo, uint64_t &hi)
{
    lo = c[0] + b;
    hi = c[1] + (lo < b ? 1u : 0u);
}

// ------------------------------------------------------------------------
// Philox4x32-10: random/internal/philox_generic_4x.hpp:40-79,
// random/internal/philox_constants.hpp:52-56,82-87.  Counter words are the
// four little-endian u32 of the 128-bit counter.
// ------------------------------------------------------------------------
// three-input XOR in one gfx950 VALU op (v_bitop3_b32, truth table 0x96)
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Each round is two v_mad_u64_u32 (hi and lo of the 32x32 product in one op)
// and two v_bitop3_b32; round keys are scalar adds.
__device__ __forceinline__ void philox4x32_10(uint32_t k0, uint32_t k1,
    uint32_t &s0, uint32_t &s1, uint32_t &s2, uint32_t &s3)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p2 = (uint64_t)s2 * 0xCD9E8D57u;
        const uint64_t p0 = (uint64_t)s0 * 0xD2511F53u;
        const uint32_t n0 =
            xor3(k0 + 0x9E3779B9u * (uint32_t)r, s1, (uint32_t)(p2 >> 32));
        const uint32_t n2 =
            xor3(k1 + 0xBB67AE85u * (uint32_t)r, s3, (uint32_t)(p0 >> 32));
        s1 = (uint32_t)p2;
        s3 = (uint32_t)p0;
        s0 = n0;
        s2 = n2;
    }
}

struct PhiloxGen {
    static constexpr int KIND = PHILOX;
    static constexpr int W32 = 4; // u32 words per block
    uint32_t k0, k1;
    uint64_t c0, c1;

    __device__ __forceinline__ void init(const StreamArg &a)
    {
        k0 = (uint32_t)a.key[0];
        k1 = (uint32_t)a.key[1];
        c0 = a.ctr[0];
        c1 = a.ctr[1];
    }
    __device__ __forceinline__ void prepare(uint32_t *) {}
    __device__ __forceinline__ void block(uint64_t b, uint32_t *w) const
    {
        const uint64_t lo = c0 + b;
        const uint64_t hi = c1 + (lo < b ? 1u : 0u);
        uint32_t s0 = (uint32_t)lo, s1 = (uint32_t)(lo >> 32);
        uint32_t s2 = (uint32_t)hi, s3 = (uint32_t)(hi >> 32);
        philox4x32_10(k0, k1, s0, s1, s2, s3);
        w[0] = s0;
        w[1] = s1;
        w[2] = s2;
        w[3] = s3;
    }
};

// ------------------------------------------------------------------------
// Threefry4x64-20: random/threefry.hpp:109-121 (parity, zero tweak),
// random/internal/threefry_common.hpp:212-247 (injection),
// random/internal/threefry_generic_4x64.hpp:40-108 (schedule),
// random/internal/threefry_constants.hpp:83-89 (rotations).
// ------------------------------------------------------------------------
// 64-bit add as a 32-bit add and add-with-carry (v_add_co_u32 +
// v_addc_co_u32); measured faster in the Threefry rounds than the single
// 64-bit v_lshl_add_u64 the compiler otherwise selects.
__device__ __forceinline__ uint64_t add64(uint64_t a, uint64_t b)
{
    unsigned c, c2;
    const uint32_t lo = __builtin_addc((uint32_t)a, (uint32_t)b, 0u, &c);
    const uint32_t hi =
        __builtin_addc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c, &c2);
    return ((uint64_t)hi << 32) | lo;
}

// 64-bit rotate left by a constant as two 32-bit funnel shifts
// (v_alignbit_b32); a rotation by 32 or more first swaps the halves.
template <int L>
__device__ __forceinline__ uint64_t rotl64(uint64_t x)
{
    static_assert(L > 0 && L < 64, "rotation");
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint32_t a = L >= 32 ? lo : hi;
    const uint32_t b = L >= 32 ? hi : lo;
    constexpr int R = L & 31;
    if (R == 0)
        return ((uint64_t)a << 32) | b;
    const uint32_t nh = __builtin_amdgcn_alignbit(a, b, 32 - R);
    const uint32_t nl = __builtin_amdgcn_alignbit(b, a, 32 - R);
    return ((uint64_t)nh << 32) | nl;
}

struct ThreefryGen {
    static constexpr int KIND = THREEFRY;
    static constexpr int W32 = 8;
    uint64_t k[5];
    uint64_t c[4];

    __device__ __forceinline__ void init(const StreamArg &a)
    {
        k[0] = a.key[0];
        k[1] = a.key[1];
        k[2] = a.key[2];
        k[3] = a.key[3];
        k[4] = 0x1BD11BDAA9FC1A22ull ^ k[0] ^ k[1] ^ k[2] ^ k[3];
        c[0] = a.ctr[0];
        c[1] = a.ctr[1];
        c[2] = a.ctr[2];
        c[3] = a.ctr[3];
    }
    __device__ __forceinline__ void prepare(uint32_t *) {}

    template <int S>
    __device__ __forceinline__ void inject(
        uint64_t &s0, uint64_t &s1, uint64_t &s2, uint64_t &s3) const
    {
        s0 = add64(s0, k[(S + 0) % 5]);
        s1 = add64(s1, k[(S + 1) % 5]);
        s2 = add64(s2, k[(S + 2) % 5]);
        s3 = add64(s3, k[(S + 3) % 5] + (uint64_t)S);
    }

    template <int R0, int R1>
    __device__ __forceinline__ static void round(
        uint64_t &s0, uint64_t &s1, uint64_t &s2, uint64_t &s3)
    {
        s0 = add64(s0, s1);
        s2 = add64(s2, s3);
        const uint64_t t1 = rotl64<R0>(s1);
        const uint64_t t3 = rotl64<R1>(s3);
        s1 = s2 ^ t3;
        s3 = s0 ^ t1;
    }

    template <int S>
    __device__ __forceinline__ void quad(
        uint64_t &s0, uint64_t &s1, uint64_t &s2, uint64_t &s3) const
    {
        // rounds 4S+1 .. 4S+4, rotation index (N-1) % 8
        if (S % 2 == 0) {
            round<14, 16>(s0, s1, s2, s3);
            round<52, 57>(s0, s1, s2, s3);
            round<23, 40>(s0, s1, s2, s3);
            round<5, 37>(s0, s1, s2, s3);
        } else {
            round<25, 33>(s0, s1, s2, s3);
            round<46, 12>(s0, s1, s2, s3);
            round<58, 22>(s0, s1, s2, s3);
            round<32, 32>(s0, s1, s2, s3);
        }
        inject<S + 1>(s0, s1, s2, s3);
    }

    __device__ __forceinline__ void block(uint64_t b, uint32_t *w) const
    {
        uint64_t s0 = c[0] + b;
        uint64_t carry = s0 < b ? 1u : 0u;
        uint64_t s1 = c[1] + carry;
        carry = (carry && s1 == 0) ? 1u : 0u;
        uint64_t s2 = c[2] + carry;
        carry = (carry && s2 == 0) ? 1u : 0u;
        uint64_t s3 = c[3] + carry;
        inject<0>(s0, s1, s2, s3);
        quad<0>(s0, s1, s2, s3);
        quad<1>(s0, s1, s2, s3);
        quad<2>(s0, s1, s2, s3);
        quad<3>(s0, s1, s2, s3);
        quad<4>(s0, s1, s2, s3);
        w[0] = (uint32_t)s0;
        w[1] = (uint32_t)(s0 >> 32);
        w[2] = (uint32_t)s1;
        w[3] = (uint32_t)(s1 >> 32);
        w[4] = (uint32_t)s2;
        w[5] = (uint32_t)(s2 >> 32);
        w[6] = (uint32_t)s3;
        w[7] = (uint32_t)(s3 >> 32);
    }
};

// ------------------------------------------------------------------------
// ARS-5 with software AES rounds: random/internal/aes_generic.hpp:254-306
// (round keys K + N*Weyl per 64-bit lane), :413-491 (T-table round and last
// round), random/internal/aes_constants.hpp:41 (T1..T3 are byte rotations
// of T0; the S-box is byte 1 of T0).  T0 lives in LDS replicated 32 times,
// one copy per bank, so lane l reads copy (l % 32) and a wavefront's
// lookups never conflict.
// ------------------------------------------------------------------------
constexpr int ARS_REPL = 32;
constexpr int ARS_LDS_WORDS = 256 * ARS_REPL;

extern __constant__ uint32_t c_aes_t0[256];

struct ArsGen {
    static constexpr int KIND = ARS;
    static constexpr int W32 = 4;
    uint64_t K0, K1;
    uint64_t c0, c1;
    const uint32_t *tab; // LDS, lane-offset applied

    __device__ __forceinline__ void init(const StreamArg &a)
    {
        K0 = a.key[0] | (a.key[1] << 32);
        K1 = a.key[2] | (a.key[3] << 32);
        c0 = a.ctr[0];
        c1 = a.ctr[1];
    }

    // Whole workgroup: fill the replicated table, then point at this lane's
    // copy.  Must be reached by every thread of the workgroup.
    __device__ __forceinline__ void prepare(uint32_t *lds)
    {
        for (int i = threadIdx.x; i < ARS_LDS_WORDS; i += blockDim.x)
            lds[i] = c_aes_t0[i / ARS_REPL];
        __syncthreads();
        tab = lds + (threadIdx.x % ARS_REPL);
    }

    __device__ __forceinline__ uint32_t T(uint32_t x) const
    {
        return tab[x * ARS_REPL];
    }

    __device__ __forceinline__ static uint32_t rotl(uint32_t x, int l)
    {
        return __builtin_amdgcn_alignbit(x, x, 32 - l);
    }

    __device__ __forceinline__ void block(uint64_t b, uint32_t *w) const
    {
        const uint64_t lo = c0 + b;
        const uint64_t hi = c1 + (lo < b ? 1u : 0u);
        uint32_t s[4] = {(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi,
            (uint32_t)(hi >> 32)};
        s[0] ^= (uint32_t)K0;
        s[1] ^= (uint32_t)(K0 >> 32);
        s[2] ^= (uint32_t)K1;
        s[3] ^= (uint32_t)(K1 >> 32);
#pragma unroll
        for (int r = 1; r < 5; ++r) {
            const uint64_t k0 = K0 + 0x9E3779B97F4A7C15ull * (uint64_t)r;
            const uint64_t k1 = K1 + 0xBB67AE8584CAA73Bull * (uint64_t)r;
            const uint32_t rk[4] = {(uint32_t)k0, (uint32_t)(k0 >> 32),
                (uint32_t)k1, (uint32_t)(k1 >> 32)};
            uint32_t t[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                t[j] = xor3(xor3(rk[j], T(s[j] & 0xFF),
                                rotl(T((s[(j + 1) & 3] >> 8) & 0xFF), 8)),
                    rotl(T((s[(j + 2) & 3] >> 16) & 0xFF), 16),
                    rotl(T(s[(j + 3) & 3] >> 24), 24));
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                s[j] = t[j];
        }
        const uint64_t k0 = K0 + 0x9E3779B97F4A7C15ull * 5u;
        const uint64_t k1 = K1 + 0xBB67AE8584CAA73Bull * 5u;
        const uint32_t rk[4] = {(uint32_t)k0, (uint32_t)(k0 >> 32),
            (uint32_t)k1, (uint32_t)(k1 >> 32)};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            // S-box byte is byte 1 of T0; place it at byte 0..3
            const uint32_t a = T(s[j] & 0xFF);
            const uint32_t bb = T((s[(j + 1) & 3] >> 8) & 0xFF);
            const uint32_t c = T((s[(j + 2) & 3] >> 16) & 0xFF);
            const uint32_t d = T(s[(j + 3) & 3] >> 24);
            w[j] = rk[j] ^ ((a >> 8) & 0xFFu) ^ (bb & 0xFF00u) ^
                (c & 0xFF0000u) ^ ((d << 16) & 0xFF000000u);
        }
    }
};

// Correctly rounded square roots.  (HIP's __fsqrt_rn is the approximate
// hardware square root unless OCML's rounded operations are enabled; the
// builtins lower to the correctly rounded sequences under the default
// -fhip-fp32-correctly-rounded-divide-sqrt.)
__device__ __forceinline__ float sqrt_rn(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sqrt_rn(double x) { return __builtin_sqrt(x); }

// ------------------------------------------------------------------------
// U01 conversions, exact fixed point: random/internal/u01_generic.hpp:48-164
// (CC/CO/OC/OO) and :166-198 (Canonical, two u32 per double).
// ------------------------------------------------------------------------
__device__ __forceinline__ double u01_f64(int iv, uint32_t u)
{
    const double p32 = 0x1p-32, p33 = 0x1p-33;
    switch (iv) {
        case 0:
            return __dmul_rn(__dadd_rn((double)(u & 1u), (double)u), p32);
        case 1:
            return __dmul_rn((double)u, p32);
        case 2:
            return __dadd_rn(__dmul_rn((double)u, p32), p32);
        default:
            return __dadd_rn(__dmul_rn((double)u, p32), p33);
    }
}

__device__ __forceinline__ float u01_f32(int iv, uint32_t u)
{
    switch (iv) {
        case 0: {
            const uint32_t v = (u << 1) >> 7;
            return __fmul_rn((float)((v & 1u) + v), 0x1p-25f);
        }
        case 1:
            return __fmul_rn((float)(u >> 8), 0x1p-24f);
        case 2:
            return __fadd_rn(__fmul_rn((float)(u >> 8), 0x1p-24f), 0x1p-24f);
        default:
            return __fadd_rn(__fmul_rn((float)(u >> 9), 0x1p-23f), 0x1p-24f);
    }
}

__device__ __forceinline__ double u01_canonical_f64(uint32_t u0, uint32_t u1)
{
    return __dadd_rn(__dmul_rn((double)u0, 0x1p-64),
        __dadd_rn(__dmul_rn((double)u1, 0x1p-32), 0.0));
}

__device__ __forceinline__ float u01_canonical_f32(uint32_t u)
{
    return __fadd_rn(__fmul_rn((float)u, 0x1p-32f), 0.0f);
}

// ------------------------------------------------------------------------
// The same from one u64 word: 64-bit-result engines take one u64 per value
// for every real type (U01UIntType, random/u01_distribution.hpp:105-113),
// with W = 64 in u01_generic.hpp:48-164 (all exact) and Canonical
// (:166-198) one rounding of the 64-bit integer.
// ------------------------------------------------------------------------
__device__ __forceinline__ double u01_f64_u64(int iv, uint64_t u)
{
    switch (iv) {
        case 0: {
            const uint64_t v = (u << 1) >> 10; // 54 bits
            return __dmul_rn((double)((v & 1u) + v), 0x1p-54);
        }
        case 1:
            return __dmul_rn((double)(u >> 11), 0x1p-53);
        case 2:
            return __dadd_rn(__dmul_rn((double)(u >> 11), 0x1p-53), 0x1p-53);
        default:
            return __dadd_rn(__dmul_rn((double)(u >> 12), 0x1p-52), 0x1p-53);
    }
}

__device__ __forceinline__ float u01_f32_u64(int iv, uint64_t u)
{
    switch (iv) {
        case 0: {
            const uint32_t v = (uint32_t)((u << 1) >> 39); // 25 bits
            return __fmul_rn((float)((v & 1u) + v), 0x1p-25f);
        }
        case 1:
            return __fmul_rn((float)(uint32_t)(u >> 40), 0x1p-24f);
        case 2:
            return __fadd_rn(
                __fmul_rn((float)(uint32_t)(u >> 40), 0x1p-24f), 0x1p-24f);
        default:
            return __fadd_rn(
                __fmul_rn((float)(uint32_t)(u >> 41), 0x1p-23f), 0x1p-24f);
    }
}

// u rounded to the nearest double (ties to even): the two halves are exact
// in double and their sum rounds once
__device__ __forceinline__ double u64_to_f64_rn(uint64_t u)
{
    return __dadd_rn(__dmul_rn((double)(uint32_t)(u >> 32), 0x1p32),
        (double)(uint32_t)u);
}

// u rounded to the nearest float (ties to even), by integer rounding of the
// top 24 bits
__device__ __forceinline__ float u64_to_f32_rn(uint64_t u)
{
    if (u < (1ull << 24))
        return (float)(uint32_t)u;
    const int lz = __clzll((long long)u);
    const uint64_t v = u << lz; // leading one at bit 63
    uint32_t m = (uint32_t)(v >> 40);
    const uint64_t rest = v & ((1ull << 40) - 1);
    const uint64_t half = 1ull << 39;
    if (rest > half || (rest == half && (m & 1u)))
        ++m; // may reach 2^24, still exact
    return __builtin_ldexpf((float)m, 40 - lz);
}

__device__ __forceinline__ double u01_canonical_f64_u64(uint64_t u)
{
    return __dmul_rn(u64_to_f64_rn(u), 0x1p-64);
}

__device__ __forceinline__ float u01_canonical_f32_u64(uint64_t u)
{
    return __fmul_rn(u64_to_f32_rn(u), 0x1p-64f);
}

} // namespace mckl_hip
