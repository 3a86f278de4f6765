// rng.hip -- bulk fills of the counter-based engines on gfx950: raw words,
// fused U01 conversions and chunked Box-Muller Normal, plus the host-side
// engine state machine that mirrors CounterEngine (random/counter.hpp).
//
// Layout: each lane computes whole counter blocks (16 or 32 bytes) and
// stores them with 16-byte vector stores, so a wavefront writes 1-2 KiB of
// contiguous HBM per store instruction.  A call that starts inside the
// engine's buffered block is handled by indexing in "block space": block 0
// is the buffered block, and the first `skip` elements are masked.
#include "common.hpp"
#include "crmath.hpp"
#include "engines.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace mckl_hip {

std::string &last_error_ref()
{
    static thread_local std::string s;
    return s;
}

__constant__ uint32_t c_aes_t0[256];
// tables of the correctly rounded double log/sin/cos (crmath.hpp)
__device__ crm::Tables d_crm_tables;

// ---------------------------------------------------------------------------
// AES T0 table from GF(2^8) arithmetic (FIPS-197 5.1.1); layout as
// random/internal/aes_constants.hpp:41 (bytes 2S, S, S, 3S).
// ---------------------------------------------------------------------------
static uint8_t gf_mul(uint8_t a, uint8_t b)
{
    uint8_t p = 0;
    for (int i = 0; i < 8; ++i) {
        if (b & 1)
            p ^= a;
        const uint8_t hi = a & 0x80;
        a = (uint8_t)(a << 1);
        if (hi)
            a ^= 0x1B;
        b >>= 1;
    }
    return p;
}

static void build_t0(uint32_t *t0)
{
    for (int x = 0; x < 256; ++x) {
        uint8_t inv = 0;
        for (int y = 1; x && y < 256; ++y)
            if (gf_mul((uint8_t)x, (uint8_t)y) == 1) {
                inv = (uint8_t)y;
                break;
            }
        const uint8_t b = inv;
        const uint8_t s = (uint8_t)(b ^ ((b << 1) | (b >> 7)) ^
            ((b << 2) | (b >> 6)) ^ ((b << 3) | (b >> 5)) ^
            ((b << 4) | (b >> 4)) ^ 0x63);
        t0[x] = (uint32_t)gf_mul(s, 2) | ((uint32_t)s << 8) |
            ((uint32_t)s << 16) | ((uint32_t)gf_mul(s, 3) << 24);
    }
}

int ensure_device_tables()
{
    static std::mutex mu;
    static bool done[64] = {false};
    int dev = 0;
    MCKL_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        return set_error(MCKL_HIP_EINVAL, "device index %d unsupported", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!done[dev]) {
        uint32_t t0[256];
        build_t0(t0);
        MCKL_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_aes_t0), t0, sizeof(t0)));
        static crm::Tables crm_host;
        static bool crm_built = false;
        if (!crm_built) {
            crm::build_tables(&crm_host);
            crm_built = true;
        }
        MCKL_HIP_TRY(hipMemcpyToSymbol(
            HIP_SYMBOL(d_crm_tables), &crm_host, sizeof(crm_host)));
        done[dev] = true;
    }
    return MCKL_HIP_OK;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
constexpr int kMaxBatch = 16; // streams per launch (kernel-argument payload)
constexpr int kBlock = 256;
constexpr int kUnroll = 2; // blocks per lane per iteration (16-byte path)

struct FillArg {
    StreamArg s;
    uint64_t skip;   // elements of block space before the call's first one
    uint32_t vec_ok; // (out - skip) is 16-byte aligned
    uint32_t pad;
};

struct FillBatch {
    FillArg a[kMaxBatch];
};

template <class Gen>
struct LdsWords {
    static constexpr int value = 1;
};
template <>
struct LdsWords<ArsGen> {
    static constexpr int value = ARS_LDS_WORDS;
};

struct ConvU32 {
    using T = uint32_t;
    static constexpr int Q = 1;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return w[0];
    }
};
struct ConvU64 {
    using T = uint64_t;
    static constexpr int Q = 2;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    }
};
struct ConvU01F64 {
    using T = double;
    static constexpr int Q = 1;
    int iv;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return u01_f64(iv, w[0]);
    }
};
struct ConvU01F32 {
    using T = float;
    static constexpr int Q = 1;
    int iv;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return u01_f32(iv, w[0]);
    }
};
// UniformReal: U01CO then muladd(u, b - a, a) without FMA, two roundings
// (random/uniform_real_distribution.hpp:48-58, math/vmf.hpp:697-722)
struct ConvUniformF64 {
    using T = double;
    static constexpr int Q = 1;
    double a, d; // d = b - a
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return __dadd_rn(__dmul_rn(u01_f64(MCKL_HIP_U01_CO, w[0]), d), a);
    }
};
struct ConvUniformF32 {
    using T = float;
    static constexpr int Q = 1;
    float a, d;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return __fadd_rn(__fmul_rn(u01_f32(MCKL_HIP_U01_CO, w[0]), d), a);
    }
};
struct ConvCanF64 {
    using T = double;
    static constexpr int Q = 2;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return u01_canonical_f64(w[0], w[1]);
    }
};
struct ConvCanF32 {
    using T = float;
    static constexpr int Q = 1;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return u01_canonical_f32(w[0]);
    }
};

// One u64 word per value (64-bit-result engines): two u32 of the block,
// little-endian, as the engine's result words are laid out
__device__ __forceinline__ uint64_t word64(const uint32_t *w)
{
    return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}
struct ConvU01F64W {
    using T = double;
    static constexpr int Q = 2;
    int iv;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return u01_f64_u64(iv, word64(w));
    }
};
struct ConvU01F32W {
    using T = float;
    static constexpr int Q = 2;
    int iv;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return u01_f32_u64(iv, word64(w));
    }
};
struct ConvCanF64W {
    using T = double;
    static constexpr int Q = 2;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return u01_canonical_f64_u64(word64(w));
    }
};
struct ConvCanF32W {
    using T = float;
    static constexpr int Q = 2;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return u01_canonical_f32_u64(word64(w));
    }
};
struct ConvUniformF64W {
    using T = double;
    static constexpr int Q = 2;
    double a, d;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return __dadd_rn(__dmul_rn(u01_f64_u64(MCKL_HIP_U01_CO, word64(w)), d),
            a);
    }
};
struct ConvUniformF32W {
    using T = float;
    static constexpr int Q = 2;
    float a, d;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return __fadd_rn(__fmul_rn(u01_f32_u64(MCKL_HIP_U01_CO, word64(w)), d),
            a);
    }
};

// UniformBits narrower than the result word: its low bits
struct ConvLow16 {
    using T = uint16_t;
    static constexpr int Q = 1;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return (uint16_t)w[0];
    }
};
struct ConvLow16W {
    using T = uint16_t;
    static constexpr int Q = 2;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return (uint16_t)w[0];
    }
};
struct ConvLow32W {
    using T = uint32_t;
    static constexpr int Q = 2;
    __device__ __forceinline__ T operator()(const uint32_t *w) const
    {
        return w[0];
    }
};

// 16-byte vector stores through L2 (plain, not non-temporal): with 32 or 64
// bytes per lane a store instruction covers only part of each cache line,
// and L2 must merge the pieces of a line before it goes to HBM; measured on
// MI355X the non-temporal form of this pattern runs at 2 TB/s, the plain
// form at 6 TB/s.
template <typename T, int N>
__device__ __forceinline__ void store_vec(T *p, const T (&v)[N])
{
    static_assert((sizeof(T) * N) % 16 == 0, "vector store granularity");
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    constexpr int NV = (int)(sizeof(T) * N / 16);
    const u32x4 *src = reinterpret_cast<const u32x4 *>(v);
    u32x4 *dst = reinterpret_cast<u32x4 *>(p);
#pragma unroll
    for (int i = 0; i < NV; ++i)
        dst[i] = src[i];
}

// Element-granular fill.  Block b of stream y covers elements
// [b*EPB - skip, (b+1)*EPB - skip) of that stream's output.
template <class Gen, class Conv>
__global__ __launch_bounds__(kBlock) void fill_kernel(FillBatch batch,
    Conv conv, uint64_t nblocks, uint64_t n, typename Conv::T *out)
{
    using T = typename Conv::T;
    constexpr int EPB = Gen::W32 / Conv::Q;
    __shared__ uint32_t lds[LdsWords<Gen>::value];
    const FillArg &arg = batch.a[blockIdx.y];
    Gen g;
    g.init(arg.s);
    g.prepare(lds);
    T *o = out + (uint64_t)blockIdx.y * n;
    const int64_t skip = (int64_t)arg.skip;
    const bool vec_ok = arg.vec_ok != 0;
    // first block holding an output element: lanes start there, so a call
    // that begins on a block boundary stores lane-aligned
    const uint64_t bf = (uint64_t)skip / EPB;
    if constexpr (sizeof(T) * EPB == 32) {
        // 32 output bytes per block.  A wave covers 64 consecutive blocks
        // (2 KiB of output) per iteration; lane pair (2j, 2j+1) computes
        // blocks j and 32 + j and the pair swaps one 16-byte half, so that
        // each of the two stores writes 1 KiB lane-contiguously.
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const int lane = threadIdx.x & 63;
        const int odd = lane & 1;
        // the wave index is uniform: keep it (and the loop bounds) scalar
        const uint64_t wave = (uint64_t)blockIdx.x * (kBlock / 64) +
            (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / 64);
        for (uint64_t wb = bf + wave * 64; wb < nblocks; wb += nwaves * 64) {
            const uint64_t b = wb + (uint64_t)(odd * 32 + (lane >> 1));
            uint32_t w[Gen::W32];
            g.block(b, w); // arithmetic only: harmless past nblocks
            T v[EPB];
#pragma unroll
            for (int j = 0; j < EPB; ++j)
                v[j] = conv(w + j * Conv::Q);
            const int64_t ew = (int64_t)(wb * EPB) - skip;
            if (vec_ok && ew >= 0 && ew + 64 * EPB <= (int64_t)n) {
                // wave-uniform branch: every lane takes part in the swap
                const u32x4 *src = reinterpret_cast<const u32x4 *>(v);
                const u32x4 own = odd ? src[1] : src[0];
                const u32x4 give = odd ? src[0] : src[1];
                u32x4 recv;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    recv[c] = (uint32_t)__builtin_amdgcn_mov_dpp(
                        (int)give[c], 0xB1, 0xF, 0xF, false); // quad_perm 1032
                u32x4 *dst = reinterpret_cast<u32x4 *>(o + ew);
                dst[lane] = odd ? recv : own;
                dst[64 + lane] = odd ? own : recv;
            } else if (b < nblocks) {
                const int64_t e0 = (int64_t)(b * EPB) - skip;
#pragma unroll
                for (int j = 0; j < EPB; ++j) {
                    const int64_t e = e0 + j;
                    if (e >= 0 && e < (int64_t)n)
                        o[e] = v[j];
                }
            }
        }
        return;
    }
    // 16 output bytes per block (or 64 for Threefry doubles): each lane
    // takes kUnroll blocks kBlock apart per iteration, so a workgroup's
    // stores stay lane-contiguous and two blocks are in flight per lane
    const uint64_t stride = (uint64_t)gridDim.x * kBlock * kUnroll;
    for (uint64_t b0 = bf + (uint64_t)blockIdx.x * kBlock * kUnroll +
             threadIdx.x;
         b0 < nblocks; b0 += stride) {
        T v[kUnroll][EPB];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            uint32_t w[Gen::W32];
            g.block(b0 + (uint64_t)u * kBlock, w);
#pragma unroll
            for (int j = 0; j < EPB; ++j)
                v[u][j] = conv(w + j * Conv::Q);
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const uint64_t b = b0 + (uint64_t)u * kBlock;
            if (b >= nblocks)
                break;
            const int64_t e0 = (int64_t)(b * EPB) - skip;
            // a block of 8 output bytes (floats from 64-bit words) is stored
            // element by element
            if constexpr ((sizeof(T) * EPB) % 16 == 0) {
                if (vec_ok && e0 >= 0 && e0 + EPB <= (int64_t)n) {
                    store_vec(o + e0, v[u]);
                    continue;
                }
            }
#pragma unroll
            for (int j = 0; j < EPB; ++j) {
                const int64_t e = e0 + j;
                if (e >= 0 && e < (int64_t)n)
                    o[e] = v[u][j];
            }
        }
    }
}

// One element per lane, for conversions whose elements straddle blocks
// (Canonical<double> at an odd word offset).  Element e uses words
// skipw + 2e, skipw + 2e + 1 of block space.
template <class Gen>
__global__ __launch_bounds__(kBlock) void fill_canonical_straddle_kernel(
    FillBatch batch, uint64_t n, double *out)
{
    __shared__ uint32_t lds[LdsWords<Gen>::value];
    const FillArg &arg = batch.a[blockIdx.y];
    Gen g;
    g.init(arg.s);
    g.prepare(lds);
    double *o = out + (uint64_t)blockIdx.y * n;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += stride) {
        const uint64_t p = arg.skip + 2 * e; // word position in block space
        uint32_t w0[Gen::W32], w1[Gen::W32];
        g.block(p / Gen::W32, w0);
        g.block(p / Gen::W32 + 1, w1);
        const uint32_t j = (uint32_t)(p % Gen::W32);
        const uint32_t a = w0[j];
        const uint32_t b = j + 1 < (uint32_t)Gen::W32 ? w0[j + 1] : w1[0];
        o[e] = u01_canonical_f64(a, b);
    }
}

// ---- Normal (chunked Box-Muller) ------------------------------------------
template <typename T>
struct NormalTraits;
template <>
struct NormalTraits<float> {
    static constexpr int K = 2048; // BufferSize<float>
};
template <>
struct NormalTraits<double> {
    static constexpr int K = 1024; // BufferSize<double>
};

// U01OC of a value's result word(s): one u32, or one u64 (Q = 2)
template <typename T, int Q>
__device__ __forceinline__ T u01oc(const uint32_t *w);
template <>
__device__ __forceinline__ float u01oc<float, 1>(const uint32_t *w)
{
    return u01_f32(MCKL_HIP_U01_OC, w[0]);
}
template <>
__device__ __forceinline__ double u01oc<double, 1>(const uint32_t *w)
{
    return u01_f64(MCKL_HIP_U01_OC, w[0]);
}
template <>
__device__ __forceinline__ float u01oc<float, 2>(const uint32_t *w)
{
    return u01_f32_u64(MCKL_HIP_U01_OC, word64(w));
}
template <>
__device__ __forceinline__ double u01oc<double, 2>(const uint32_t *w)
{
    return u01_f64_u64(MCKL_HIP_U01_OC, word64(w));
}

// Box-Muller on one (s, theta) pair of U01OC values with the reference's op
// order (normal_distribution.hpp:58-77).  Float transcendentals are
// evaluated in double and rounded once (the correctly rounded value).
__device__ __forceinline__ void box_muller(float u1, float u2,
    float mean, float sd, float &o_sin, float &o_cos)
{
    float s = (float)log((double)u1);
    s = __fmul_rn(s, -2.0f);
    s = sqrt_rn(s);
    const float th = __fmul_rn(6.283185307179586476925286766559005768394f, u2);
    double sn, cs;
    sincos((double)th, &sn, &cs);
    const float a = (float)sn, b = (float)cs;
    if (sd != 1.0f)
        s = __fmul_rn(sd, s);
    if (mean != 0.0f) {
        o_sin = __fadd_rn(__fmul_rn(s, a), mean);
        o_cos = __fadd_rn(__fmul_rn(s, b), mean);
    } else {
        o_sin = __fmul_rn(s, a);
        o_cos = __fmul_rn(s, b);
    }
}

// Double: log, sin and cos correctly rounded (crmath.hpp), so a result can
// differ from the reference's only where its libm misrounds.
__device__ __forceinline__ void box_muller(double u1, double u2,
    double mean, double sd, double &o_sin, double &o_cos)
{
    double s = crm::log_cr(u1, d_crm_tables.log_c);
    s = __dmul_rn(s, -2.0);
    s = sqrt_rn(s);
    const double th = __dmul_rn(6.283185307179586, u2);
    double a, b;
    crm::sincos_cr(th, d_crm_tables.sin_a, d_crm_tables.cos_a, a, b);
    if (sd != 1.0)
        s = __dmul_rn(sd, s);
    if (mean != 0.0) {
        o_sin = __dadd_rn(__dmul_rn(s, a), mean);
        o_cos = __dadd_rn(__dmul_rn(s, b), mean);
    } else {
        o_sin = __dmul_rn(s, a);
        o_cos = __dmul_rn(s, b);
    }
}

// odd tail: normal_distribution.hpp:93-100
__device__ __forceinline__ float box_muller_tail(
    float u, float v, float mean, float sd)
{
    const float l = (float)log((double)u);
    const float r = sqrt_rn(__fmul_rn(-2.0f, l));
    const float c = (float)cos(
        (double)__fmul_rn(6.283185307179586476925286766559005768394f, v));
    return __fadd_rn(mean, __fmul_rn(__fmul_rn(sd, r), c));
}

__device__ __forceinline__ double box_muller_tail(
    double u, double v, double mean, double sd)
{
    const double r =
        sqrt_rn(__dmul_rn(-2.0, crm::log_cr(u, d_crm_tables.log_c)));
    double sn, c;
    crm::sincos_cr(__dmul_rn(6.283185307179586, v), d_crm_tables.sin_a,
        d_crm_tables.cos_a, sn, c);
    return __dadd_rn(mean, __dmul_rn(__dmul_rn(sd, r), c));
}

// Word p (relative to the call's first word) of a stream whose first word is
// word `skip` of block 0.
template <class Gen>
__device__ __forceinline__ uint32_t word_at(const Gen &g, uint64_t skip,
    uint64_t p)
{
    const uint64_t q = p + skip;
    uint32_t w[Gen::W32];
    g.block(q / Gen::W32, w);
    uint32_t r = w[0];
#pragma unroll
    for (int j = 1; j < Gen::W32; ++j)
        if ((uint32_t)(q % Gen::W32) == (uint32_t)j)
            r = w[j];
    return r;
}

// The result words of value p (relative to the call's first value) of a
// stream whose first u32 word is word `skip` of block 0: Q u32 words
template <class Gen, int Q>
__device__ __forceinline__ void value_words(const Gen &g, uint64_t skip,
    uint64_t p, uint32_t (&w)[Q])
{
#pragma unroll
    for (int q = 0; q < Q; ++q)
        w[q] = word_at(g, skip, p * Q + q);
}

// Full chunks of K values.  Q u32 words per value (1, or 2 for one u64 word
// of a 64-bit-result engine), V = W32 / Q values per block.  Task (c, g):
// pairs g*V .. g*V+V-1 of chunk c.  With skip == 0 the s-words and
// theta-words of a task are exactly one block each; otherwise each window
// straddles two blocks (skip counts u32 words).  Straddle instantiates the
// second form; a launch whose streams all have skip == 0 uses the first,
// which then holds no runtime-indexed word arrays (those live in scratch).
template <class Gen, typename T, int Q, bool Straddle>
__global__ __launch_bounds__(kBlock) void normal_kernel(FillBatch batch,
    uint64_t nchunks, uint64_t n, T mean, T sd, T *out)
{
    constexpr int K = NormalTraits<T>::K;
    constexpr int H = K / 2;
    constexpr int W = Gen::W32;
    constexpr int V = W / Q;
    constexpr int G = H / V; // tasks per chunk
    __shared__ uint32_t lds[LdsWords<Gen>::value];
    const FillArg &arg = batch.a[blockIdx.y];
    Gen g;
    g.init(arg.s);
    g.prepare(lds);
    T *o = out + (uint64_t)blockIdx.y * n;
    const uint32_t skip = (uint32_t)arg.skip;
    const uint64_t ntask = nchunks * G;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
         t < ntask; t += stride) {
        const uint64_t c = t / G;
        const uint32_t gi = (uint32_t)(t % G);
        const uint64_t bs = c * (K / V) + gi;  // block of first s-word
        const uint64_t bt = bs + H / V;        // block of first theta-word
        uint32_t ws[W], wt[W];
        if constexpr (!Straddle) {
            g.block(bs, ws);
            g.block(bt, wt);
        } else {
            uint32_t a[W], b[W];
            g.block(bs, a);
            g.block(bs + 1, b);
#pragma unroll
            for (int j = 0; j < W; ++j)
                ws[j] = j + skip < (uint32_t)W ? a[j + skip] : b[j + skip - W];
            g.block(bt, a);
            g.block(bt + 1, b);
#pragma unroll
            for (int j = 0; j < W; ++j)
                wt[j] = j + skip < (uint32_t)W ? a[j + skip] : b[j + skip - W];
        }
        T vs[V], vc[V];
#pragma unroll
        for (int j = 0; j < V; ++j)
            box_muller(u01oc<T, Q>(ws + j * Q), u01oc<T, Q>(wt + j * Q), mean,
                sd, vs[j], vc[j]);
        T *os = o + c * K + (uint64_t)gi * V;
        if constexpr ((sizeof(T) * V) % 16 == 0) {
            if (arg.vec_ok) {
                store_vec(os, vs);
                store_vec(os + H, vc);
                continue;
            }
        }
#pragma unroll
        for (int j = 0; j < V; ++j) {
            os[j] = vs[j];
            os[H + j] = vc[j];
        }
    }
}

// Last partial chunk (L = n % K values, nu = L/2 pairs) and the odd tail
// (two scalar U01OC draws, one result word each).
template <class Gen, typename T, int Q>
__global__ __launch_bounds__(kBlock) void normal_tail_kernel(FillBatch batch,
    uint64_t nchunks, uint64_t n, T mean, T sd, T *out)
{
    constexpr int K = NormalTraits<T>::K;
    __shared__ uint32_t lds[LdsWords<Gen>::value];
    const FillArg &arg = batch.a[blockIdx.y];
    Gen g;
    g.init(arg.s);
    g.prepare(lds);
    T *o = out + (uint64_t)blockIdx.y * n;
    const uint64_t base = nchunks * K;
    const uint64_t L = n - base;
    const uint64_t nu = L / 2;
    for (uint64_t p = threadIdx.x; p <= nu; p += blockDim.x) {
        uint32_t wa[Q], wb[Q];
        if (p < nu) {
            value_words<Gen, Q>(g, arg.skip, base + p, wa);
            value_words<Gen, Q>(g, arg.skip, base + nu + p, wb);
            T a, b;
            box_muller(u01oc<T, Q>(wa), u01oc<T, Q>(wb), mean, sd, a, b);
            o[base + p] = a;
            o[base + nu + p] = b;
        } else if (L % 2 == 1) {
            value_words<Gen, Q>(g, arg.skip, base + L, wa);
            value_words<Gen, Q>(g, arg.skip, base + L + 1, wb);
            o[base + L - 1] =
                box_muller_tail(u01oc<T, Q>(wa), u01oc<T, Q>(wb), mean, sd);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side: engine state machine (random/counter.hpp)
// ---------------------------------------------------------------------------
static int ctr_words(int kind) { return kind == THREEFRY ? 4 : 2; }
static int block_w32(int kind) { return kind == THREEFRY ? 8 : 4; }

static int check_engine(const mckl_hip_engine *e)
{
    if (!e)
        return set_error(MCKL_HIP_EINVAL, "engine is NULL");
    if (e->kind < PHILOX || e->kind > ARS)
        return set_error(MCKL_HIP_EINVAL, "unknown engine kind %d", e->kind);
    if (e->rbits != 32 && e->rbits != 64)
        return set_error(MCKL_HIP_EINVAL, "rbits must be 32 or 64");
    const uint32_t M = (uint32_t)(block_w32(e->kind) * 32 / e->rbits);
    if (e->index > M)
        return set_error(MCKL_HIP_EINVAL, "engine index %u > %u", e->index, M);
    return MCKL_HIP_OK;
}

static uint32_t words_per_block(const mckl_hip_engine *e)
{
    return (uint32_t)(block_w32(e->kind) * 32 / e->rbits);
}

// ++ctr with ripple carry: random/internal/increment_generic.hpp:73-103
static void ctr_inc(uint64_t *c, int k)
{
    for (int i = 0; i < k; ++i)
        if (++c[i] != 0)
            return;
}

// ctr += v with full carry (the effect of v single increments)
static void ctr_add(uint64_t *c, int k, uint64_t v)
{
    const uint64_t old = c[0];
    c[0] += v;
    if (c[0] < old)
        ctr_inc(c + 1, k - 1);
}

// increment(ctr, nskip): random/internal/increment_generic.hpp:208-217,
// including its strict-less carry test (used only by discard).
static void ctr_increment_skip(uint64_t *c, int k, uint64_t nskip)
{
    if (c[0] < UINT64_MAX - nskip) {
        c[0] += nskip;
    } else {
        c[0] += nskip;
        ctr_inc(c + 1, k - 1);
    }
}

// State transition of one bulk call operator()(n, r): counter.hpp:142-165
static void bulk_call(mckl_hip_engine *e, uint64_t n)
{
    const uint32_t M = words_per_block(e);
    const uint64_t remain = M - e->index;
    if (n <= remain) {
        e->index += (uint32_t)n;
        return;
    }
    n -= remain;
    const uint64_t m = n / M;
    ctr_add(e->ctr, ctr_words(e->kind), m);
    ctr_inc(e->ctr, ctr_words(e->kind));
    e->index = (uint32_t)(n - m * M);
}

// `count` bulk calls of k words each, k a multiple of M and k > M: after the
// first call index is fixed and each further call adds k/M to the counter.
static void bulk_calls(mckl_hip_engine *e, uint64_t k, uint64_t count)
{
    if (count == 0)
        return;
    bulk_call(e, k);
    if (count > 1) {
        const uint32_t M = words_per_block(e);
        ctr_add(e->ctr, ctr_words(e->kind), (count - 1) * (k / M));
    }
}

// scalar operator()(): counter.hpp:134-142
static void scalar_call(mckl_hip_engine *e)
{
    const uint32_t M = words_per_block(e);
    if (e->index == M) {
        ctr_inc(e->ctr, ctr_words(e->kind));
        e->index = 0;
    }
    e->index++;
}

// Position of the next word: counter of the block holding it, and its u32
// offset within that block.
static void next_position(const mckl_hip_engine *e, StreamArg &s, uint64_t &skip32)
{
    const uint32_t M = words_per_block(e);
    std::memcpy(s.ctr, e->ctr, sizeof(s.ctr));
    std::memcpy(s.key, e->key, sizeof(s.key));
    if (e->index < M) {
        skip32 = (uint64_t)e->index * (e->rbits / 32);
    } else {
        ctr_inc(s.ctr, ctr_words(e->kind));
        skip32 = 0;
    }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
// Workgroups per launch: streaming kernels want a deep grid (65536 groups of
// 256 measured best for the fills); ARS amortises a 32 KiB LDS table per
// workgroup, so it runs a smaller grid with more blocks per lane.
static int grid_cap(int kind) { return kind == ARS ? 2048 : 65536; }

template <class Gen, class Conv>
static int launch_fill(const std::vector<FillArg> &args, Conv conv,
    uint64_t nblocks, uint64_t n, typename Conv::T *out, hipStream_t st)
{
    for (size_t b0 = 0; b0 < args.size(); b0 += kMaxBatch) {
        FillBatch batch;
        std::memset(&batch, 0, sizeof(batch));
        const size_t cnt = std::min<size_t>(kMaxBatch, args.size() - b0);
        for (size_t i = 0; i < cnt; ++i)
            batch.a[i] = args[b0 + i];
        constexpr uint64_t EPB = Gen::W32 / Conv::Q;
        uint64_t bf = nblocks;
        for (size_t i = 0; i < cnt; ++i)
            bf = std::min<uint64_t>(bf, batch.a[i].skip / EPB);
        const uint64_t per_group = (sizeof(typename Conv::T) * EPB == 32)
            ? (uint64_t)kBlock
            : (uint64_t)kBlock * kUnroll;
        const uint64_t want = ceil_div(nblocks - bf, per_group);
        const uint64_t cap = (uint64_t)grid_cap(Gen::KIND) /
            std::max<size_t>(1, cnt);
        const uint32_t gx = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(want, std::max<uint64_t>(1, cap)));
        dim3 grid(gx, (uint32_t)cnt);
        fill_kernel<Gen, Conv><<<grid, kBlock, 0, st>>>(
            batch, conv, nblocks, n, out + b0 * n);
        MCKL_HIP_CHECK_LAUNCH("fill_kernel");
    }
    return MCKL_HIP_OK;
}

template <class Gen>
static int launch_straddle(const std::vector<FillArg> &args, uint64_t n,
    double *out, hipStream_t st)
{
    for (size_t b0 = 0; b0 < args.size(); b0 += kMaxBatch) {
        FillBatch batch;
        std::memset(&batch, 0, sizeof(batch));
        const size_t cnt = std::min<size_t>(kMaxBatch, args.size() - b0);
        for (size_t i = 0; i < cnt; ++i)
            batch.a[i] = args[b0 + i];
        const uint32_t gx = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(ceil_div(n, kBlock), 4096));
        fill_canonical_straddle_kernel<Gen>
            <<<dim3(gx, (uint32_t)cnt), kBlock, 0, st>>>(batch, n, out + b0 * n);
        MCKL_HIP_CHECK_LAUNCH("fill_canonical_straddle_kernel");
    }
    return MCKL_HIP_OK;
}

template <class Gen, typename T>
static int launch_normal(const std::vector<FillArg> &args, int Q, uint64_t n,
    T mean, T sd, T *out, hipStream_t st)
{
    constexpr int K = NormalTraits<T>::K;
    const uint64_t nchunks = n / K;
    const uint64_t ntask = nchunks * (K / 2 / (Gen::W32 / Q));
    for (size_t b0 = 0; b0 < args.size(); b0 += kMaxBatch) {
        FillBatch batch;
        std::memset(&batch, 0, sizeof(batch));
        const size_t cnt = std::min<size_t>(kMaxBatch, args.size() - b0);
        for (size_t i = 0; i < cnt; ++i)
            batch.a[i] = args[b0 + i];
        if (ntask) {
            const uint64_t cap = std::max<uint64_t>(1,
                (uint64_t)grid_cap(Gen::KIND) /
                    std::max<size_t>(1, cnt));
            const uint32_t gx = (uint32_t)std::max<uint64_t>(
                1, std::min<uint64_t>(ceil_div(ntask, kBlock), cap));
            bool straddle = false;
            for (size_t i = 0; i < cnt; ++i)
                straddle = straddle || batch.a[i].skip != 0;
            const dim3 grid(gx, (uint32_t)cnt);
            T *o = out + b0 * n;
            if (Q == 2 && straddle)
                normal_kernel<Gen, T, 2, true><<<grid, kBlock, 0, st>>>(
                    batch, nchunks, n, mean, sd, o);
            else if (Q == 2)
                normal_kernel<Gen, T, 2, false><<<grid, kBlock, 0, st>>>(
                    batch, nchunks, n, mean, sd, o);
            else if (straddle)
                normal_kernel<Gen, T, 1, true><<<grid, kBlock, 0, st>>>(
                    batch, nchunks, n, mean, sd, o);
            else
                normal_kernel<Gen, T, 1, false><<<grid, kBlock, 0, st>>>(
                    batch, nchunks, n, mean, sd, o);
            MCKL_HIP_CHECK_LAUNCH("normal_kernel");
        }
        if (n % K != 0) {
            if (Q == 2)
                normal_tail_kernel<Gen, T, 2>
                    <<<dim3(1, (uint32_t)cnt), kBlock, 0, st>>>(
                        batch, nchunks, n, mean, sd, out + b0 * n);
            else
                normal_tail_kernel<Gen, T, 1>
                    <<<dim3(1, (uint32_t)cnt), kBlock, 0, st>>>(
                        batch, nchunks, n, mean, sd, out + b0 * n);
            MCKL_HIP_CHECK_LAUNCH("normal_tail_kernel");
        }
    }
    return MCKL_HIP_OK;
}

// dispatch on engine kind
template <template <class> class Op, typename... A>
static int by_kind(int kind, A &&...a)
{
    switch (kind) {
        case PHILOX:
            return Op<PhiloxGen>::run(std::forward<A>(a)...);
        case THREEFRY:
            return Op<ThreefryGen>::run(std::forward<A>(a)...);
        case ARS:
            return Op<ArsGen>::run(std::forward<A>(a)...);
    }
    return set_error(MCKL_HIP_EINVAL, "unknown engine kind %d", kind);
}

template <class Gen>
struct FillOp {
    template <class Conv>
    static int run(const std::vector<FillArg> &args, Conv conv,
        uint64_t nblocks, uint64_t n, typename Conv::T *out, hipStream_t st)
    {
        return launch_fill<Gen, Conv>(args, conv, nblocks, n, out, st);
    }
};

template <class Gen>
struct StraddleOp {
    static int run(const std::vector<FillArg> &args, uint64_t n, double *out,
        hipStream_t st)
    {
        return launch_straddle<Gen>(args, n, out, st);
    }
};

template <class Gen>
struct NormalOp {
    template <typename T>
    static int run(const std::vector<FillArg> &args, int Q, uint64_t n, T mean,
        T sd, T *out, hipStream_t st)
    {
        return launch_normal<Gen, T>(args, Q, n, mean, sd, out, st);
    }
};

static int common_checks(const mckl_hip_engine *e, size_t nstreams)
{
    for (size_t s = 0; s < nstreams; ++s) {
        int rc = check_engine(e + s);
        if (rc)
            return rc;
        if (e[s].kind != e[0].kind || e[s].rbits != e[0].rbits)
            return set_error(MCKL_HIP_EINVAL,
                "all engines of a multi-stream call must share kind and width");
    }
    return MCKL_HIP_OK;
}

// Element fill over nstreams engines with Q u32 words per element (Q in
// {1,2}), advancing each engine by the reference's chunked call sequence
// (chunk words `kwords`, 0 = a single bulk call of n*Q result words).
template <class Conv>
static int element_fill(mckl_hip_engine *e, size_t nstreams, uint64_t n,
    typename Conv::T *out, Conv conv, uint64_t kwords, hipStream_t st)
{
    int rc = common_checks(e, nstreams);
    if (rc)
        return rc;
    if (n == 0 || nstreams == 0)
        return MCKL_HIP_OK;
    if (!out)
        return set_error(MCKL_HIP_EINVAL, "output pointer is NULL");
    rc = ensure_device_tables();
    if (rc)
        return rc;
    using T = typename Conv::T;
    const int W32 = block_w32(e[0].kind);
    std::vector<FillArg> args(nstreams);
    uint64_t nblocks = 0;
    bool straddle = false;
    for (size_t s = 0; s < nstreams; ++s) {
        uint64_t skip32 = 0;
        next_position(e + s, args[s].s, skip32);
        if (skip32 % Conv::Q != 0)
            straddle = true;
        args[s].skip = skip32 / Conv::Q;
        const uintptr_t base = reinterpret_cast<uintptr_t>(out + s * n) -
            args[s].skip * sizeof(T);
        args[s].vec_ok = (base % 16 == 0) ? 1u : 0u;
        args[s].pad = 0;
        nblocks = std::max<uint64_t>(nblocks,
            ceil_div(skip32 + n * Conv::Q, (uint64_t)W32));
    }
    if (straddle) {
        // only Canonical<double> (Q = 2) at an odd word offset gets here
        for (size_t s = 0; s < nstreams; ++s) {
            uint64_t skip32 = 0;
            next_position(e + s, args[s].s, skip32);
            args[s].skip = skip32;
        }
        rc = by_kind<StraddleOp>(e[0].kind, args, n,
            reinterpret_cast<double *>(out), st);
    } else {
        rc = by_kind<FillOp>(e[0].kind, args, conv, nblocks, n, out, st);
    }
    if (rc)
        return rc;
    const uint64_t rwords = n * Conv::Q * 32 / (uint64_t)e[0].rbits;
    for (size_t s = 0; s < nstreams; ++s) {
        if (kwords == 0) {
            bulk_call(e + s, rwords);
        } else {
            bulk_calls(e + s, kwords, rwords / kwords);
            bulk_call(e + s, rwords % kwords);
        }
    }
    return MCKL_HIP_OK;
}

template <typename T>
static int normal_fill(mckl_hip_engine *e, size_t nstreams, uint64_t n, T mean,
    T sd, T *out, hipStream_t st)
{
    int rc = common_checks(e, nstreams);
    if (rc)
        return rc;
    if (!(sd > 0))
        return set_error(MCKL_HIP_EINVAL,
            "NormalDistribution requires stddev > 0"); // normal_distribution.hpp:45-48
    if (n == 0 || nstreams == 0)
        return MCKL_HIP_OK;
    if (!out)
        return set_error(MCKL_HIP_EINVAL, "output pointer is NULL");
    rc = ensure_device_tables();
    if (rc)
        return rc;
    constexpr int K = NormalTraits<T>::K;
    std::vector<FillArg> args(nstreams);
    for (size_t s = 0; s < nstreams; ++s) {
        uint64_t skip32 = 0;
        next_position(e + s, args[s].s, skip32);
        args[s].skip = skip32;
        const uintptr_t base = reinterpret_cast<uintptr_t>(out + s * n);
        args[s].vec_ok = (base % 16 == 0) ? 1u : 0u;
        args[s].pad = 0;
    }
    rc = by_kind<NormalOp>(e[0].kind, args, e[0].rbits / 32, n, mean, sd,
        out, st);
    if (rc)
        return rc;
    for (size_t s = 0; s < nstreams; ++s) {
        bulk_calls(e + s, K, n / K);
        bulk_call(e + s, n % K);
        if (n % 2 == 1) {
            scalar_call(e + s);
            scalar_call(e + s);
        }
    }
    return MCKL_HIP_OK;
}

// raw generator kernel: n blocks for counters ctr+1..ctr+n
static int raw_kernel(int kind, uint64_t *ctr, size_t n, void *d_r,
    const uint64_t *key, void *stream)
{
    if (!ctr || !key)
        return set_error(MCKL_HIP_EINVAL, "ctr/key is NULL");
    if (n == 0)
        return MCKL_HIP_OK;
    if (!d_r)
        return set_error(MCKL_HIP_EINVAL, "output pointer is NULL");
    int rc = ensure_device_tables();
    if (rc)
        return rc;
    std::vector<FillArg> args(1);
    std::memset(args.data(), 0, sizeof(FillArg));
    std::memcpy(args[0].s.ctr, ctr, sizeof(uint64_t) * ctr_words(kind));
    std::memcpy(args[0].s.key, key, sizeof(uint64_t) * 4);
    ctr_inc(args[0].s.ctr, ctr_words(kind));
    args[0].skip = 0;
    args[0].vec_ok = (reinterpret_cast<uintptr_t>(d_r) % 16 == 0) ? 1u : 0u;
    const int W32 = block_w32(kind);
    rc = by_kind<FillOp>(kind, args, ConvU32(), (uint64_t)n, (uint64_t)n * W32,
        static_cast<uint32_t *>(d_r), as_stream(stream));
    if (rc)
        return rc;
    ctr_add(ctr, ctr_words(kind), n);
    return MCKL_HIP_OK;
}

// one block on the device, copied back (synchronous)
__global__ void one_block_kernel(int kind, StreamArg a, uint32_t *out)
{
    __shared__ uint32_t lds[ARS_LDS_WORDS];
    uint32_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (kind == PHILOX) {
        PhiloxGen g;
        g.init(a);
        if (threadIdx.x == 0)
            g.block(0, w);
    } else if (kind == THREEFRY) {
        ThreefryGen g;
        g.init(a);
        if (threadIdx.x == 0)
            g.block(0, w);
    } else {
        ArsGen g;
        g.init(a);
        g.prepare(lds);
        if (threadIdx.x == 0)
            g.block(0, w);
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < 8; ++j)
            out[j] = w[j];
}

static int block_on_device(int kind, const uint64_t *ctr, const uint64_t *key,
    uint32_t *h_out)
{
    int rc = ensure_device_tables();
    if (rc)
        return rc;
    StreamArg a;
    std::memset(&a, 0, sizeof(a));
    std::memcpy(a.ctr, ctr, sizeof(a.ctr));
    std::memcpy(a.key, key, sizeof(a.key));
    uint32_t *d = nullptr;
    MCKL_HIP_TRY(hipMalloc(&d, 32));
    one_block_kernel<<<1, kBlock>>>(kind, a, d);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess)
        err = hipMemcpy(h_out, d, 32, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return hip_status(err, "one_block_kernel");
}

// U01Canonical<double> fill used by the resampling sequences
// (u01_distribution.hpp:115-129: two u32 per double, 1024 per bulk call)
int rand_canonical_f64(mckl_hip_engine *e, size_t n, double *d_r, hipStream_t st)
{
    return element_fill(e, 1, n, d_r, ConvCanF64(), 2048, st);
}

} // namespace mckl_hip

using namespace mckl_hip;

extern "C" {

const char *mckl_hip_last_error(void) { return last_error_ref().c_str(); }

int mckl_hip_device_math(
    const double **d_log_c, const double **d_sin_a, const double **d_cos_a)
{
    if (!d_log_c || !d_sin_a || !d_cos_a)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    int rc = ensure_device_tables();
    if (rc)
        return rc;
    crm::Tables *t = nullptr;
    MCKL_HIP_TRY(hipGetSymbolAddress(
        reinterpret_cast<void **>(&t), HIP_SYMBOL(d_crm_tables)));
    *d_log_c = t->log_c;
    *d_sin_a = t->sin_a;
    *d_cos_a = t->cos_a;
    return MCKL_HIP_OK;
}

int mckl_hip_engine_words(const mckl_hip_engine *e)
{
    int rc = check_engine(e);
    return rc ? rc : (int)words_per_block(e);
}

int mckl_hip_engine_init(
    mckl_hip_engine *e, int kind, int rbits, const uint64_t *key)
{
    if (!e || !key)
        return set_error(MCKL_HIP_EINVAL, "engine/key is NULL");
    std::memset(e, 0, sizeof(*e));
    e->kind = kind;
    e->rbits = rbits;
    int rc = check_engine(e);
    if (rc)
        return rc;
    // key words: Philox4x32 two u32, ARS four u32, Threefry four u64
    const int nk = kind == PHILOX ? 2 : 4;
    for (int i = 0; i < nk; ++i)
        e->key[i] = kind == THREEFRY ? key[i] : (uint32_t)key[i];
    e->index = words_per_block(e);
    return MCKL_HIP_OK;
}

int mckl_hip_engine_seed(mckl_hip_engine *e, uint64_t s)
{
    int rc = check_engine(e);
    if (rc)
        return rc;
    // seed(result_type s): counter.hpp:97-103
    const uint64_t sv = e->rbits == 32 ? (uint32_t)s : s;
    uint64_t key[4] = {sv, 0, 0, 0};
    return mckl_hip_engine_init(e, e->kind, e->rbits, key);
}

int mckl_hip_engine_set_ctr(mckl_hip_engine *e, const uint64_t *ctr)
{
    int rc = check_engine(e);
    if (rc)
        return rc;
    if (!ctr)
        return set_error(MCKL_HIP_EINVAL, "ctr is NULL");
    std::memset(e->ctr, 0, sizeof(e->ctr));
    std::memcpy(e->ctr, ctr, sizeof(uint64_t) * ctr_words(e->kind));
    e->index = words_per_block(e);
    return MCKL_HIP_OK;
}

int mckl_hip_engine_discard(mckl_hip_engine *e, uint64_t nskip)
{
    int rc = check_engine(e);
    if (rc)
        return rc;
    if (nskip == 0)
        return MCKL_HIP_OK;
    const uint32_t M = words_per_block(e);
    const uint64_t remain = M - e->index;
    if (nskip <= remain) {
        e->index += (uint32_t)nskip;
        return MCKL_HIP_OK;
    }
    nskip -= remain;
    e->index = M;
    ctr_increment_skip(e->ctr, ctr_words(e->kind), nskip / M);
    ctr_inc(e->ctr, ctr_words(e->kind));
    e->index = (uint32_t)(nskip % M);
    return MCKL_HIP_OK;
}

int mckl_hip_engine_block(const mckl_hip_engine *e, void *h_block)
{
    int rc = check_engine(e);
    if (rc)
        return rc;
    if (!h_block)
        return set_error(MCKL_HIP_EINVAL, "output is NULL");
    uint32_t w[8];
    rc = block_on_device(e->kind, e->ctr, e->key, w);
    if (rc)
        return rc;
    std::memcpy(h_block, w, (size_t)block_w32(e->kind) * 4);
    return MCKL_HIP_OK;
}

int mckl_hip_engine_next(mckl_hip_engine *e, uint64_t *h_out)
{
    int rc = check_engine(e);
    if (rc)
        return rc;
    if (!h_out)
        return set_error(MCKL_HIP_EINVAL, "output is NULL");
    scalar_call(e);
    uint32_t w[8];
    rc = block_on_device(e->kind, e->ctr, e->key, w);
    if (rc)
        return rc;
    const uint32_t j = e->index - 1;
    *h_out = e->rbits == 32
        ? (uint64_t)w[j]
        : ((uint64_t)w[2 * j] | ((uint64_t)w[2 * j + 1] << 32));
    return MCKL_HIP_OK;
}

int mckl_hip_philox4x32_kernel(
    uint64_t *ctr, size_t n, void *d_r, const uint32_t *key, void *stream)
{
    if (!key)
        return set_error(MCKL_HIP_EINVAL, "key is NULL");
    uint64_t k[4] = {key[0], key[1], 0, 0};
    return raw_kernel(PHILOX, ctr, n, d_r, k, stream);
}

int mckl_hip_threefry4x64_kernel(
    uint64_t *ctr, size_t n, void *d_r, const uint64_t *key, void *stream)
{
    return raw_kernel(THREEFRY, ctr, n, d_r, key, stream);
}

int mckl_hip_ars_kernel(
    uint64_t *ctr, size_t n, void *d_r, const uint32_t *key, void *stream)
{
    if (!key)
        return set_error(MCKL_HIP_EINVAL, "key is NULL");
    uint64_t k[4] = {key[0], key[1], key[2], key[3]};
    return raw_kernel(ARS, ctr, n, d_r, k, stream);
}

int mckl_hip_rand(mckl_hip_engine *e, size_t n, void *d_r, void *stream)
{
    int rc = check_engine(e);
    if (rc)
        return rc;
    if (e->rbits == 32)
        return element_fill(e, 1, n, static_cast<uint32_t *>(d_r), ConvU32(), 0,
            as_stream(stream));
    return element_fill(
        e, 1, n, static_cast<uint64_t *>(d_r), ConvU64(), 0, as_stream(stream));
}

int mckl_hip_rand_uniform_bits(
    mckl_hip_engine *e, int ubits, size_t n, void *d_r, void *stream)
{
    int rc = check_engine(e);
    if (rc)
        return rc;
    if (ubits != 16 && ubits != 32 && ubits != 64)
        return set_error(MCKL_HIP_EINVAL,
            "UniformBits needs 16, 32 or 64 bits, not %d", ubits);
    // at least as wide as the result: the result words, reinterpreted
    // (uniform_bits_distribution.hpp:72-80)
    if (ubits >= e->rbits)
        return mckl_hip_rand(e, n * (size_t)(ubits / e->rbits), d_r, stream);
    // narrower: one result word per value, chunks of BufferSize<UIntType>
    // = 8192 / sizeof(UIntType) result words (:82-97)
    const uint64_t k = 8192 / (uint64_t)(ubits / 8);
    hipStream_t st = as_stream(stream);
    if (ubits == 16 && e->rbits == 32)
        return element_fill(e, 1, n, static_cast<uint16_t *>(d_r),
            ConvLow16(), k, st);
    if (ubits == 16)
        return element_fill(e, 1, n, static_cast<uint16_t *>(d_r),
            ConvLow16W(), k, st);
    return element_fill(e, 1, n, static_cast<uint32_t *>(d_r), ConvLow32W(),
        k, st);
}

static int u01_f64_impl(mckl_hip_engine *e, size_t ns, int interval, size_t n,
    double *d_r, void *stream)
{
    if (interval < 0 || interval > 4)
        return set_error(MCKL_HIP_EINVAL, "unknown U01 interval %d", interval);
    if (!e)
        return set_error(MCKL_HIP_EINVAL, "engine is NULL");
    // BufferSize<double> = 1024 elements per bulk call (common.hpp:69-86);
    // 64-bit-result engines take one u64 word per value, Canonical included
    if (e[0].rbits == 64) {
        if (interval == MCKL_HIP_U01_CANONICAL)
            return element_fill(e, ns, n, d_r, ConvCanF64W(), 1024,
                as_stream(stream));
        ConvU01F64W c;
        c.iv = interval;
        return element_fill(e, ns, n, d_r, c, 1024, as_stream(stream));
    }
    if (interval == MCKL_HIP_U01_CANONICAL)
        return element_fill(e, ns, n, d_r, ConvCanF64(), 2048, as_stream(stream));
    ConvU01F64 c;
    c.iv = interval;
    return element_fill(e, ns, n, d_r, c, 1024, as_stream(stream));
}

int mckl_hip_rand_u01_f64(
    mckl_hip_engine *e, int interval, size_t n, double *d_r, void *stream)
{
    return u01_f64_impl(e, 1, interval, n, d_r, stream);
}

int mckl_hip_rand_u01_f64_streams(mckl_hip_engine *e, size_t nstreams,
    int interval, size_t n, double *d_r, void *stream)
{
    return u01_f64_impl(e, nstreams, interval, n, d_r, stream);
}

int mckl_hip_rand_u01_f32(
    mckl_hip_engine *e, int interval, size_t n, float *d_r, void *stream)
{
    if (interval < 0 || interval > 4)
        return set_error(MCKL_HIP_EINVAL, "unknown U01 interval %d", interval);
    if (e && e->rbits == 64) {
        if (interval == MCKL_HIP_U01_CANONICAL)
            return element_fill(e, 1, n, d_r, ConvCanF32W(), 2048,
                as_stream(stream));
        ConvU01F32W c;
        c.iv = interval;
        return element_fill(e, 1, n, d_r, c, 2048, as_stream(stream));
    }
    if (interval == MCKL_HIP_U01_CANONICAL)
        return element_fill(e, 1, n, d_r, ConvCanF32(), 2048, as_stream(stream));
    ConvU01F32 c;
    c.iv = interval;
    return element_fill(e, 1, n, d_r, c, 2048, as_stream(stream));
}

int mckl_hip_rand_uniform_real_f64(mckl_hip_engine *e, size_t n, double a,
    double b, double *d_r, void *stream)
{
    if (!(a < b))
        return set_error(MCKL_HIP_EINVAL,
            "**UniformRealDistribution** constructed with invalid arguments");
    if (e && e->rbits == 64) {
        ConvUniformF64W c;
        c.a = a;
        c.d = b - a;
        return element_fill(e, 1, n, d_r, c, 1024, as_stream(stream));
    }
    ConvUniformF64 c;
    c.a = a;
    c.d = b - a;
    return element_fill(e, 1, n, d_r, c, 1024, as_stream(stream));
}

int mckl_hip_rand_uniform_real_f32(mckl_hip_engine *e, size_t n, float a,
    float b, float *d_r, void *stream)
{
    if (!(a < b))
        return set_error(MCKL_HIP_EINVAL,
            "**UniformRealDistribution** constructed with invalid arguments");
    if (e && e->rbits == 64) {
        ConvUniformF32W c;
        c.a = a;
        c.d = b - a;
        return element_fill(e, 1, n, d_r, c, 2048, as_stream(stream));
    }
    ConvUniformF32 c;
    c.a = a;
    c.d = b - a;
    return element_fill(e, 1, n, d_r, c, 2048, as_stream(stream));
}

int mckl_hip_rand_normal_f32(mckl_hip_engine *e, size_t n, float mean,
    float sd, float *d_r, void *stream)
{
    return normal_fill<float>(e, 1, n, mean, sd, d_r, as_stream(stream));
}

int mckl_hip_rand_normal_f64(mckl_hip_engine *e, size_t n, double mean,
    double sd, double *d_r, void *stream)
{
    return normal_fill<double>(e, 1, n, mean, sd, d_r, as_stream(stream));
}

int mckl_hip_rand_normal_f32_streams(mckl_hip_engine *e, size_t nstreams,
    size_t n, float mean, float sd, float *d_r, void *stream)
{
    return normal_fill<float>(e, nstreams, n, mean, sd, d_r, as_stream(stream));
}

int mckl_hip_rand_normal_f64_streams(mckl_hip_engine *e, size_t nstreams,
    size_t n, double mean, double sd, double *d_r, void *stream)
{
    return normal_fill<double>(e, nstreams, n, mean, sd, d_r, as_stream(stream));
}

} // extern "C"

// ---------------------------------------------------------------------------
// SeedGenerator keys and fresh per-engine streams (RNGSetVector)
// random/seed.hpp:54-341, random/rng_set.hpp:81-120
// ---------------------------------------------------------------------------
namespace mckl_hip {

// Threefry2x32-20 / Threefry2x64-20 of one block under key 0 with zero tweak
// (random/internal/threefry_common.hpp:212-247: word 0 += par[S % 3],
// word 1 += par[(S + 1) % 3] + S; round s0 += s1, s1 = rotl(s1, R) ^ s0;
// rotations random/internal/threefry_constants.hpp:62-80).
__device__ __forceinline__ void threefry2x32_key0(uint32_t &s0, uint32_t &s1)
{
    constexpr int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
    const uint32_t par[3] = {0u, 0u, 0x1BD11BDAu};
#pragma unroll
    for (int S = 0; S <= 5; ++S) {
        s0 += par[S % 3];
        s1 += par[(S + 1) % 3] + (uint32_t)S;
        if (S == 5)
            break;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            s0 += s1;
            const int l = R[(S * 4 + q) % 8];
            s1 = __builtin_amdgcn_alignbit(s1, s1, 32 - l) ^ s0;
        }
    }
}

__device__ __forceinline__ void threefry2x64_key0(uint64_t &s0, uint64_t &s1)
{
    const uint64_t par[3] = {0ull, 0ull, 0x1BD11BDAA9FC1A22ull};
#define MCKL_TF2_ROUND(L)                                                     \
    s0 = add64(s0, s1);                                                       \
    s1 = rotl64<L>(s1) ^ s0;
#pragma unroll
    for (int S = 0; S <= 5; ++S) {
        s0 = add64(s0, par[S % 3]);
        s1 = add64(s1, par[(S + 1) % 3] + (uint64_t)S);
        if (S == 5)
            break;
        if (S % 2 == 0) {
            MCKL_TF2_ROUND(16)
            MCKL_TF2_ROUND(42)
            MCKL_TF2_ROUND(12)
            MCKL_TF2_ROUND(31)
        } else {
            MCKL_TF2_ROUND(16)
            MCKL_TF2_ROUND(32)
            MCKL_TF2_ROUND(24)
            MCKL_TF2_ROUND(21)
        }
    }
#undef MCKL_TF2_ROUND
}

// n consecutive get() from internal seed `seed`; keys in engine layout
__global__ __launch_bounds__(kBlock) void k_seed_keys(int kind, uint64_t seed,
    uint64_t np, uint64_t rank, uint64_t maxs, int randomize, uint64_t n,
    uint64_t *keys)
{
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n)
        return;
    uint64_t s = seed + i;
    uint64_t k[4] = {0, 0, 0, 0};
    if (kind == PHILOX) {
        s = (s % maxs) * np + rank;
        uint32_t w0 = (uint32_t)s, w1 = (uint32_t)(s >> 32);
        if (randomize)
            threefry2x32_key0(w0, w1);
        k[0] = w0;
        k[1] = w1;
    } else if (kind == ARS) {
        uint64_t w0 = s, w1 = rank;
        if (randomize)
            threefry2x64_key0(w0, w1);
        k[0] = (uint32_t)w0;
        k[1] = w0 >> 32;
        k[2] = (uint32_t)w1;
        k[3] = w1 >> 32;
    } else {
        k[0] = s;
        k[3] = rank;
        if (randomize) {
            StreamArg a{};
            a.ctr[0] = s;
            a.ctr[3] = rank;
            ThreefryGen g;
            g.init(a); // key 0
            uint32_t w[8];
            g.block(0, w);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                k[j] = (uint64_t)w[2 * j] | ((uint64_t)w[2 * j + 1] << 32);
        }
    }
    uint64_t *o = keys + 4 * i;
    o[0] = k[0];
    o[1] = k[1];
    o[2] = k[2];
    o[3] = k[3];
}

// Engine e (key keys[4e..4e+3], counter 0, nothing buffered) produces `per`
// elements: stream words from block ctr + 1 on, one lane per (engine,
// block).  out[e * per + j] is element j of engine e.
template <class Gen, class Conv>
__global__ __launch_bounds__(kBlock) void k_rngset_fill(const uint64_t *keys,
    uint64_t nengines, uint64_t per, uint64_t bpe, Conv conv,
    typename Conv::T *out)
{
    using T = typename Conv::T;
    constexpr int EPB = Gen::W32 / Conv::Q;
    __shared__ uint32_t lds[LdsWords<Gen>::value];
    const uint64_t total = nengines * bpe;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    Gen g;
    StreamArg a{};
    g.init(a);
    g.prepare(lds);
    for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < total;
         t += stride) {
        const uint64_t e = t / bpe, b = t % bpe;
        a.key[0] = keys[4 * e];
        a.key[1] = keys[4 * e + 1];
        a.key[2] = keys[4 * e + 2];
        a.key[3] = keys[4 * e + 3];
        Gen ge = g;
        ge.init(a);
        uint32_t w[Gen::W32];
        ge.block(b + 1, w);
        T *o = out + e * per;
#pragma unroll
        for (int j = 0; j < EPB; ++j) {
            const uint64_t k = b * EPB + j;
            if (k < per)
                o[k] = conv(w + j * Conv::Q);
        }
    }
}

template <class Gen>
struct RngSetOp {
    template <class Conv>
    static int run(const uint64_t *keys, uint64_t nengines, uint64_t per,
        Conv conv, typename Conv::T *out, hipStream_t st)
    {
        constexpr int EPB = Gen::W32 / Conv::Q;
        const uint64_t bpe = ceil_div(per, (uint64_t)EPB);
        const uint64_t total = nengines * bpe;
        const uint64_t cap = (uint64_t)grid_cap(Gen::KIND);
        const uint32_t gx = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(ceil_div(total, kBlock), cap));
        k_rngset_fill<Gen, Conv><<<gx, kBlock, 0, st>>>(
            keys, nengines, per, bpe, conv, out);
        MCKL_HIP_CHECK_LAUNCH("k_rngset_fill");
        return MCKL_HIP_OK;
    }
};

template <class Conv>
static int rngset_fill(int kind, size_t nengines, const uint64_t *d_keys,
    size_t per, typename Conv::T *out, Conv conv, hipStream_t st)
{
    if (kind != PHILOX && kind != THREEFRY && kind != ARS)
        return set_error(MCKL_HIP_EINVAL, "unknown engine kind %d", kind);
    if (nengines == 0 || per == 0)
        return MCKL_HIP_OK;
    if (!d_keys || !out)
        return set_error(MCKL_HIP_EINVAL, "NULL device pointer");
    int rc = ensure_device_tables();
    if (rc)
        return rc;
    return by_kind<RngSetOp>(kind, d_keys, (uint64_t)nengines, (uint64_t)per,
        conv, out, st);
}

} // namespace mckl_hip

extern "C" {

int mckl_hip_seed_keys(int kind, uint64_t seed, uint64_t np, uint64_t rank,
    int randomize, size_t n, uint64_t *d_keys, void *stream)
{
    if (kind != PHILOX && kind != THREEFRY && kind != ARS)
        return set_error(MCKL_HIP_EINVAL, "unknown engine kind %d", kind);
    if (np == 0 || rank >= np)
        return set_error(MCKL_HIP_EINVAL,
            "partition needs np > 0 and rank < np"); // seed.hpp:173-184
    if (n == 0)
        return MCKL_HIP_OK;
    if (!d_keys)
        return set_error(MCKL_HIP_EINVAL, "NULL device pointer");
    const uint64_t maxs = np > 1 ? (UINT64_MAX - rank) / np + 1 : UINT64_MAX;
    const uint64_t grid = ceil_div((uint64_t)n, (uint64_t)kBlock);
    k_seed_keys<<<(uint32_t)grid, kBlock, 0, as_stream(stream)>>>(kind, seed,
        np, rank, maxs, randomize, (uint64_t)n, d_keys);
    MCKL_HIP_CHECK_LAUNCH("k_seed_keys");
    return MCKL_HIP_OK;
}

int mckl_hip_rngset_rand(int kind, size_t nengines, const uint64_t *d_keys,
    size_t per, uint32_t *d_r, void *stream)
{
    return rngset_fill(kind, nengines, d_keys, per, d_r, ConvU32(),
        as_stream(stream));
}

int mckl_hip_rngset_u01_f64(int kind, int interval, size_t nengines,
    const uint64_t *d_keys, size_t per, double *d_r, void *stream)
{
    if (interval < MCKL_HIP_U01_CC || interval > MCKL_HIP_U01_OO)
        return set_error(MCKL_HIP_EINVAL, "interval %d", interval);
    ConvU01F64 c;
    c.iv = interval;
    return rngset_fill(kind, nengines, d_keys, per, d_r, c, as_stream(stream));
}

int mckl_hip_rngset_u01_f32(int kind, int interval, size_t nengines,
    const uint64_t *d_keys, size_t per, float *d_r, void *stream)
{
    if (interval < MCKL_HIP_U01_CC || interval > MCKL_HIP_U01_OO)
        return set_error(MCKL_HIP_EINVAL, "interval %d", interval);
    ConvU01F32 c;
    c.iv = interval;
    return rngset_fill(kind, nengines, d_keys, per, d_r, c, as_stream(stream));
}

} // extern "C"
