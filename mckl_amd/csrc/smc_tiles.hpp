// smc_tiles.hpp -- per-tile stages of one SMC step, shared by the plain
// (one workgroup per tile, many launches) and the batched (one workgroup per
// filter, one launch) samplers.  Each function is executed by one group of
// 256 threads; `gt` is the thread index within the group and `lds` points to
// that group's scratch.  A tile is particles [tile*2048, tile*2048+2048),
// 8 contiguous particles per thread.
//
// Step t (t = 1..T) of the sampler, following SMCSampler::do_iterate
// (algorithm/smc.hpp:389-406):
//   move     x_i = a x_{p(i)} + s eps(i,t), lw_i = carry_i + log g(y_t | x_i)
//            p(i) = idx_i and carry_i = 0 after a resampling step, else
//            p(i) = i and carry_i = lw_i - m_{t-1}            -> tile max
//   sums     e_i = exp_f(lw_i - m_t), S = sum e, S2 = sum e^2 -> tile sums
//   finalize logZ += m_t + log S_t - log C_t, ESS = S^2/S2 (smc.hpp:397)
//   resample (ESS < threshold N) W_i = e_i / S, FP64 prefix sums in
//            double-double, replication counts (resample.hpp:374-407) and
//            the stable rep->index transform (resample.hpp:416-466).
#pragma once

#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "smc_index.hpp"
#include "smc_math.hpp"

namespace mckl_hip {

constexpr int kScratch = 1280; // LDS bytes per 256-thread group

struct GroupLds {
    double *d;
    float *f;
    uint32_t *u;
};

__device__ __forceinline__ GroupLds group_lds(unsigned char *base)
{
    GroupLds g;
    g.d = reinterpret_cast<double *>(base);
    g.f = reinterpret_cast<float *>(base + 64);
    g.u = reinterpret_cast<uint32_t *>(base + 128);
    return g;
}

template <typename T>
static int dev_alloc(T **p, size_t n)
{
    *p = nullptr;
    MCKL_HIP_TRY(hipMalloc(reinterpret_cast<void **>(p),
        std::max<size_t>(n, 1) * sizeof(T)));
    return MCKL_HIP_OK;
}

struct SmcArgs {
    ModelConst mc;
    uint32_t k0, k1; // particle RNG key: seed lo, seed hi
    uint32_t N;
    uint32_t ntiles;
    double tauN; // resampling threshold * N
    int scheme;  // MCKL_HIP_MULTINOMIAL, _STRATIFIED or _SYSTEMATIC
    uint32_t T;
    const float *obs; // T observation constants
};

constexpr uint32_t kZeroFlag = 0x80000000u;

__device__ __forceinline__ uint64_t tile_first(uint32_t tile, int gt)
{
    return (uint64_t)tile * kTile + (uint64_t)gt * kItems;
}

// The 8 values of a thread in a tile row: two 16-byte loads when all are
// below N (rows are padded to whole tiles and 16-byte aligned), else
// element loads with `fill` past N.
__device__ __forceinline__ void load8(const float *p, uint64_t i0, uint64_t N,
    float (&v)[kItems], float fill)
{
    if (i0 + kItems <= N) {
        const float4 a = *reinterpret_cast<const float4 *>(p + i0);
        const float4 b = *reinterpret_cast<const float4 *>(p + i0 + 4);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = a.w;
        v[4] = b.x;
        v[5] = b.y;
        v[6] = b.z;
        v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            v[k] = i0 + k < N ? p[i0 + k] : fill;
    }
}

__device__ __forceinline__ void load8(const uint32_t *p, uint64_t i0,
    uint64_t N, uint32_t (&v)[kItems], uint32_t fill)
{
    if (i0 + kItems <= N) {
        const uint4 a = *reinterpret_cast<const uint4 *>(p + i0);
        const uint4 b = *reinterpret_cast<const uint4 *>(p + i0 + 4);
        v[0] = a.x;
        v[1] = a.y;
        v[2] = a.z;
        v[3] = a.w;
        v[4] = b.x;
        v[5] = b.y;
        v[6] = b.z;
        v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            v[k] = i0 + k < N ? p[i0 + k] : fill;
    }
}

// Store the 8 values of a thread: two 16-byte stores when all are below N.
__device__ __forceinline__ void store8(uint32_t *p, uint64_t i0, uint64_t N,
    const uint32_t (&v)[kItems])
{
    if (i0 + kItems <= N) {
        *reinterpret_cast<uint4 *>(p + i0) = make_uint4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<uint4 *>(p + i0 + 4) =
            make_uint4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (i0 + k < N)
                p[i0 + k] = v[k];
    }
}

// Non-temporal form of the two 16-byte stores of a thread's 8 values, for
// outputs the next launch reads: they bypass L2 residency, so the end of
// the launch has fewer dirty lines to write back.
template <typename T>
__device__ __forceinline__ void store8_nt(T *p, uint64_t i0, uint64_t N,
    const T (&v)[kItems])
{
    static_assert(sizeof(T) == 4, "32-bit elements");
    typedef T t4 __attribute__((ext_vector_type(4)));
    if (i0 + kItems <= N) {
        const t4 a = {v[0], v[1], v[2], v[3]};
        const t4 b = {v[4], v[5], v[6], v[7]};
        __builtin_nontemporal_store(a, reinterpret_cast<t4 *>(p + i0));
        __builtin_nontemporal_store(b, reinterpret_cast<t4 *>(p + i0 + 4));
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (i0 + k < N)
                p[i0 + k] = v[k];
    }
}

// group sums of two u32 per thread; results in every thread
__device__ __forceinline__ void group_sum2_u(
    uint32_t &a, uint32_t &b, int gt, uint32_t *ldsu)
{
    a = wave_sum_u(a);
    b = wave_sum_u(b);
    if ((gt & 63) == 0) {
        ldsu[gt >> 6] = a;
        ldsu[4 + (gt >> 6)] = b;
    }
    __syncthreads();
    a = ldsu[0] + ldsu[1] + ldsu[2] + ldsu[3];
    b = ldsu[4] + ldsu[5] + ldsu[6] + ldsu[7];
    __syncthreads();
}

// ---- models ----
// A fused sampler takes its model as a value of a type M with
//   __device__ float init(float e) const         x_0 from a standard normal e
//   __device__ Step at(uint32_t t) const         the model at step t = 1..T
// where Step (any type) has
//   __device__ float move(float x, float e) const   x_t from x_{t-1} and a
//                                                  standard normal e
//   __device__ float log_weight(float x) const      log g(y_t | x_t)
// The roles of the reference's evaluation object (smp/backend_base.hpp:
// 113-183): at(t) is eval_first's per-step set-up, move/log_weight are
// eval_each on one particle.  M is passed to the kernels by value, so it
// holds plain values and device pointers.  The noise is the samplers'
// counter-based stream (smc_normal_pair: particle i, step t), the states
// FP32, the bootstrap weights lw_t = carry + log_weight(x_t).
//
// A model whose state is a vector of D > 1 floats declares
//   static constexpr int dim = D;
// and takes and returns the state through arrays of D floats:
//   __device__ void init(const float *e, float *x) const   x_0 from D normals
//   __device__ void move(const float *xp, const float *e, float *xn) const
//   __device__ float log_weight(const float *x) const
// (move and log_weight on the Step).  Component d of particle i draws the
// stream's normal of index i D + d, so a model of dimension 1 written either
// way draws the same noise.  States are stored by component (component d of
// every particle contiguous, a stride of whole tiles), so each component's
// loads and stores stay coalesced.
template <class M, class = void>
struct model_dim {
    static constexpr int value = 1;
};
template <class M>
struct model_dim<M, std::void_t<decltype(M::dim)>> {
    static constexpr int value = M::dim;
};

// The built-in linear-Gaussian and stochastic-volatility models in that
// form: ModelConst and the per-step observation constants (obs_const).
struct BuiltinModel {
    ModelConst mc;
    const float *obs; // T observation constants
    struct Step {
        ModelConst mc;
        float obs;
        __device__ __forceinline__ float move(float x, float e) const
        {
            return smc_move(mc, x, e);
        }
        __device__ __forceinline__ float log_weight(float x) const
        {
            return smc_logg(mc, obs, x);
        }
    };
    __device__ __forceinline__ float init(float e) const
    {
        return __fmul_rn(mc.sd0, e);
    }
    __device__ __forceinline__ Step at(uint32_t t) const
    {
        return Step{mc, obs[t - 1]};
    }
};

__host__ __device__ inline BuiltinModel builtin_model(const SmcArgs &A)
{
    return BuiltinModel{A.mc, A.obs};
}

// the stream's normals of indices [j0, j0 + 2 n) at step t (j0 even): the
// pairs of smc_normal_pair
template <int n2>
__device__ __forceinline__ void normals_from(
    uint32_t k0, uint32_t k1, uint64_t j0, uint64_t t, float *e)
{
#pragma unroll
    for (int q = 0; q < n2; q += 2)
        smc_normal_pair(k0, k1, (j0 + q) >> 1, t, e[q], e[q + 1]);
}

// a thread's 8 values of one state component (or of lw), plain or
// non-temporal stores, element stores past N
template <bool NT>
__device__ __forceinline__ void store8f(
    float *p, uint64_t i0, uint64_t N, const float (&v)[kItems])
{
    if constexpr (NT) {
        store8_nt(p, i0, N, v);
    } else if (i0 + kItems <= N) {
        *reinterpret_cast<float4 *>(p + i0) =
            make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(p + i0 + 4) =
            make_float4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (i0 + k < N)
                p[i0 + k] = v[k];
    }
}

// vector states (model_dim > 1): component d of particle i at x[d Np + i],
// Np = ntiles * kTile
template <class M>
__device__ __forceinline__ void tile_init_v(const SmcArgs &A, const M &model,
    uint32_t tile, float *x, float *lw, int gt)
{
    constexpr int D = model_dim<M>::value;
    const uint64_t i0 = tile_first(tile, gt);
    const uint64_t Np = (uint64_t)A.ntiles * kTile;
    float e[kItems * D];
    normals_from<kItems * D>(A.k0, A.k1, i0 * D, 0, e);
    float xs[D][kItems], zero[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        float xv[D];
        model.init(e + k * D, xv);
#pragma unroll
        for (int d = 0; d < D; ++d)
            xs[d][k] = xv[d];
        zero[k] = 0.0f;
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
        store8f<false>(x + d * Np, i0, A.N, xs[d]);
    store8f<false>(lw, i0, A.N, zero);
}

// t = 0: x_i = init(eps(i, 0)), lw_i = 0 (scalar states)
template <class M>
__device__ __forceinline__ void tile_init_s(const SmcArgs &A, const M &model,
    uint32_t tile, float *x, float *lw, int gt)
{
    const uint64_t i0 = tile_first(tile, gt);
#pragma unroll
    for (int k = 0; k < kItems; k += 2) {
        const uint64_t i = i0 + k;
        if (i >= A.N)
            break;
        float e0, e1;
        smc_normal_pair(A.k0, A.k1, i >> 1, 0, e0, e1);
        x[i] = model.init(e0);
        lw[i] = 0.0f;
        if (i + 1 < A.N) {
            x[i + 1] = model.init(e1);
            lw[i + 1] = 0.0f;
        }
    }
}

// t = 0: x_i = init(eps(i, 0)), lw_i = 0
template <class M>
__device__ __forceinline__ void tile_init(const SmcArgs &A, const M &model,
    uint32_t tile, float *x, float *lw, int gt)
{
    if constexpr (model_dim<M>::value > 1)
        tile_init_v(A, model, tile, x, lw, gt);
    else
        tile_init_s(A, model, tile, x, lw, gt);
}

// first j with E[j] > z (E nondecreasing, length n)
__device__ __forceinline__ uint32_t upper_bound_u(
    const uint32_t *E, uint32_t n, uint32_t z)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (E[mid] > z)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// Parent of slot i after the stable rep->index transform, from a code that
// is the slot itself when it kept a copy, else kZeroFlag | the slot's rank
// among the zero slots (the source of that extra copy is the first j with
// E_j > rank).  Used by the gradient sampler, which keeps these codes.
__device__ __forceinline__ uint32_t parent_of(
    uint32_t code, const uint32_t *Eincl, uint32_t N)
{
    return (code & kZeroFlag) ? upper_bound_u(Eincl, N, code & ~kZeroFlag)
                              : code;
}

// Parent of slot code a: with src == nullptr the code is the parent itself
// (idx holds resolved parents); otherwise a zero slot's code is kZeroFlag |
// its zero-slot rank r and its parent is src[r], the source of extra copy r.
// Parents are kept below N whatever the codes hold.
__device__ __forceinline__ uint32_t resolve_parent(
    uint32_t a, const uint32_t *src, uint32_t N)
{
    if (src && (a & kZeroFlag))
        a = src[a & ~kZeroFlag];
    return a < N ? a : N - 1;
}

// tile_move for vector states (model_dim > 1)
template <bool Codes, class M>
__device__ __forceinline__ float tile_move_v(const SmcArgs &A, const M &model,
    uint32_t t, uint32_t tile, const float *xin, float *xout, float *lw,
    const uint32_t *idx, const uint32_t *src, int resampled, float mprev,
    int gt, float *ldsf, float (&lout)[kItems])
{
    constexpr int D = model_dim<M>::value;
    const uint64_t i0 = tile_first(tile, gt);
    const uint64_t Np = (uint64_t)A.ntiles * kTile;
    const auto step = model.at(t);
    uint32_t par[kItems];
    float carry[kItems];
    if (resampled) {
        uint32_t a[kItems];
        load8(idx, i0, A.N, a, 0u);
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            par[k] = resolve_parent(a[k], Codes ? src : nullptr, A.N);
            carry[k] = 0.0f;
        }
    } else {
        float w[kItems];
        load8(lw, i0, A.N, w, 0.0f);
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            par[k] = i0 + k < A.N ? (uint32_t)(i0 + k) : A.N - 1;
            carry[k] = __fsub_rn(w[k], mprev);
        }
    }
    // the normals do not depend on the parents: drawn while they load
    float e[kItems * D];
    normals_from<kItems * D>(A.k0, A.k1, i0 * D, t, e);
    float xp[kItems * D];
#pragma unroll
    for (int d = 0; d < D; ++d)
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            xp[k * D + d] = xin[d * Np + par[k]];
    float xs[D][kItems];
    float lmax = -__builtin_huge_valf();
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        float xn[D];
        step.move(xp + k * D, e + k * D, xn);
        const float l = __fadd_rn(carry[k], step.log_weight(xn));
#pragma unroll
        for (int d = 0; d < D; ++d)
            xs[d][k] = xn[d];
        const bool live = i0 + k < A.N;
        lout[k] = live ? l : 0.0f;
        lmax = live ? fmaxf(lmax, l) : lmax;
    }
#pragma unroll
    for (int d = 0; d < D; ++d)
        store8f<Codes>(xout + d * Np, i0, A.N, xs[d]);
    store8f<Codes>(lw, i0, A.N, lout);
    return group_max(lmax, gt, ldsf);
}

// tile_move for scalar states
template <bool Codes, class M>
__device__ __forceinline__ float tile_move_s(const SmcArgs &A, const M &model,
    uint32_t t, uint32_t tile, const float *xin, float *xout, float *lw,
    const uint32_t *idx, const uint32_t *src, int resampled, float mprev,
    int gt, float *ldsf, float (&lout)[kItems])
{
    const uint64_t i0 = tile_first(tile, gt);
    const auto step = model.at(t);
    float lmax = -__builtin_huge_valf();
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        lout[k] = 0.0f;
    if (i0 + kItems <= A.N) {
        float xp[kItems], lo[kItems], e[kItems];
        if (resampled) {
            const uint4 a0 = *reinterpret_cast<const uint4 *>(idx + i0);
            const uint4 a1 = *reinterpret_cast<const uint4 *>(idx + i0 + 4);
            // the normals do not depend on the parents: half are generated
            // while the codes load, half while the sources load
#pragma unroll
            for (int k = 0; k < kItems / 2; k += 2)
                smc_normal_pair(A.k0, A.k1, (i0 + k) >> 1, t, e[k], e[k + 1]);
            uint32_t a[kItems] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y,
                a1.z, a1.w};
            // branch-free resolution: every slot reads src (a non-zero code
            // is the slot's own index, also below N), so the eight source
            // loads, then the eight state loads, are each in flight at once
            if constexpr (Codes) {
                uint32_t sv[kItems];
#pragma unroll
                for (int k = 0; k < kItems; ++k) {
                    const uint32_t r = a[k] & ~kZeroFlag;
                    sv[k] = src[r < A.N ? r : A.N - 1];
                }
#pragma unroll
                for (int k = kItems / 2; k < kItems; k += 2)
                    smc_normal_pair(
                        A.k0, A.k1, (i0 + k) >> 1, t, e[k], e[k + 1]);
#pragma unroll
                for (int k = 0; k < kItems; ++k)
                    a[k] = (a[k] & kZeroFlag) ? sv[k] : a[k];
            } else {
#pragma unroll
                for (int k = kItems / 2; k < kItems; k += 2)
                    smc_normal_pair(
                        A.k0, A.k1, (i0 + k) >> 1, t, e[k], e[k + 1]);
            }
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                xp[k] = xin[a[k] < A.N ? a[k] : A.N - 1];
                lo[k] = 0.0f;
            }
        } else {
            const float4 v0 = *reinterpret_cast<const float4 *>(xin + i0);
            const float4 v1 = *reinterpret_cast<const float4 *>(xin + i0 + 4);
            const float4 w0 = *reinterpret_cast<const float4 *>(lw + i0);
            const float4 w1 = *reinterpret_cast<const float4 *>(lw + i0 + 4);
            const float xs[kItems] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y,
                v1.z, v1.w};
            const float ws[kItems] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y,
                w1.z, w1.w};
#pragma unroll
            for (int k = 0; k < kItems; k += 2)
                smc_normal_pair(A.k0, A.k1, (i0 + k) >> 1, t, e[k], e[k + 1]);
#pragma unroll
            for (int k = 0; k < kItems; ++k) {
                xp[k] = xs[k];
                lo[k] = __fsub_rn(ws[k], mprev);
            }
        }
        float xn[kItems], ln[kItems];
#pragma unroll
        for (int k = 0; k < kItems; k += 2) {
            xn[k] = step.move(xp[k], e[k]);
            xn[k + 1] = step.move(xp[k + 1], e[k + 1]);
            ln[k] = __fadd_rn(lo[k], step.log_weight(xn[k]));
            ln[k + 1] = __fadd_rn(lo[k + 1], step.log_weight(xn[k + 1]));
            lmax = fmaxf(lmax, fmaxf(ln[k], ln[k + 1]));
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            lout[k] = ln[k];
        if constexpr (Codes) { // the plain sampler: read by the next launch
            store8_nt(xout, i0, A.N, xn);
            store8_nt(lw, i0, A.N, ln);
        } else {
            *reinterpret_cast<float4 *>(xout + i0) =
                make_float4(xn[0], xn[1], xn[2], xn[3]);
            *reinterpret_cast<float4 *>(xout + i0 + 4) =
                make_float4(xn[4], xn[5], xn[6], xn[7]);
            *reinterpret_cast<float4 *>(lw + i0) =
                make_float4(ln[0], ln[1], ln[2], ln[3]);
            *reinterpret_cast<float4 *>(lw + i0 + 4) =
                make_float4(ln[4], ln[5], ln[6], ln[7]);
        }
    } else {
        for (int k = 0; k < kItems; k += 2) {
            const uint64_t i = i0 + k;
            if (i >= A.N)
                break;
            float e[2];
            smc_normal_pair(A.k0, A.k1, i >> 1, t, e[0], e[1]);
            for (int q = 0; q < 2; ++q) {
                const uint64_t p = i + q;
                if (p >= A.N)
                    break;
                float xv, carry;
                if (resampled) {
                    xv = xin[resolve_parent(idx[p], Codes ? src : nullptr,
                        A.N)];
                    carry = 0.0f;
                } else {
                    xv = xin[p];
                    carry = __fsub_rn(lw[p], mprev);
                }
                const float xnew = step.move(xv, e[q]);
                const float l = __fadd_rn(carry, step.log_weight(xnew));
                xout[p] = xnew;
                lw[p] = l;
                lout[k + q] = l;
                lmax = fmaxf(lmax, l);
            }
        }
    }
    return group_max(lmax, gt, ldsf);
}

// move + weight of one tile; returns the tile max of lw (all threads) and
// the thread's new log weights in lout (0 past the end).  After a
// resampling step idx holds every slot's parent (Codes false), or its code
// resolved through src (Codes true, resolve_parent).
template <bool Codes, class M>
__device__ __forceinline__ float tile_move(const SmcArgs &A, const M &model,
    uint32_t t, uint32_t tile, const float *xin, float *xout, float *lw,
    const uint32_t *idx, const uint32_t *src, int resampled, float mprev,
    int gt, float *ldsf, float (&lout)[kItems])
{
    if constexpr (model_dim<M>::value > 1)
        return tile_move_v<Codes>(A, model, t, tile, xin, xout, lw, idx,
            src, resampled, mprev, gt, ldsf, lout);
    else
        return tile_move_s<Codes>(A, model, t, tile, xin, xout, lw, idx,
            src, resampled, mprev, gt, ldsf, lout);
}

// Weights.  With m_k the tile's max of lw and m the max over tiles, particle
// i of tile k weighs e_i = a_i r_k, a_i = exp_f(lw_i - m_k), r_k =
// exp_f(m_k - m) (tile_scale); the product of the two floats is exact in
// double.  The tile's S = sum a, S2 = sum a^2 (canonical order) and the
// double-double total of a are taken right after the move, before m is
// known; the samplers scale them by r_k (S, dd total) and r_k^2 (S2).
__device__ __forceinline__ double tile_scale(float mk, float m)
{
    return (double)smc_exp_f(__fsub_rn(mk, m));
}

// the tile's sums of a_i = exp_f(lw_i - mk) (see above); returns the
// thread's exclusive double-double prefix of a within the tile, the toff of
// tile_counts_local, bit for bit
__device__ __forceinline__ dd tile_sums_l(const SmcArgs &A, uint32_t tile,
    const float (&l)[kItems], float m, int gt, double *ldsd, double &S,
    double &S2, dd &etot)
{
    const uint64_t i0 = tile_first(tile, gt);
    double s = 0.0, s2 = 0.0;
    dd loc{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < A.N) {
            const double e = (double)smc_exp_f(__fsub_rn(l[k], m));
            s = __dadd_rn(s, e);
            s2 = __dadd_rn(s2, __dmul_rn(e, e));
            loc = dd_add(loc, dd{e, 0.0});
        }
    }
    group_sum2(s, s2, gt, ldsd);
    S = s;
    S2 = s2;
    return group_excl_scan_dd(loc, gt, ldsd, etot);
}

// tile_sums_l's S and S2 alone (same operations), for a sampler that takes
// the double-double total only on the steps that resample (tile_dd_total)
__device__ __forceinline__ void tile_sums2(const SmcArgs &A, uint32_t tile,
    const float (&l)[kItems], float m, int gt, double *ldsd, double &S,
    double &S2)
{
    const uint64_t i0 = tile_first(tile, gt);
    double s = 0.0, s2 = 0.0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < A.N) {
            const double e = (double)smc_exp_f(__fsub_rn(l[k], m));
            s = __dadd_rn(s, e);
            s2 = __dadd_rn(s2, __dmul_rn(e, e));
        }
    }
    group_sum2(s, s2, gt, ldsd);
    S = s;
    S2 = s2;
}

// tile_sums_l's double-double total etot alone (same operations)
__device__ __forceinline__ dd tile_dd_total(const SmcArgs &A, uint32_t tile,
    const float (&l)[kItems], float m, int gt, double *ldsd)
{
    const uint64_t i0 = tile_first(tile, gt);
    dd loc{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < A.N) {
            const double e = (double)smc_exp_f(__fsub_rn(l[k], m));
            loc = dd_add(loc, dd{e, 0.0});
        }
    }
    dd etot;
    (void)group_excl_scan_dd(loc, gt, ldsd, etot);
    return etot;
}

// Uniform sequences of step t (sorted, M = N):
//   systematic u_j = j*delta + u0d  (resample.hpp:122-135), u0d = u01*delta
//   stratified u_j = delta*(c_j + j) (resample.hpp:95-120)
struct UniformSeq {
    int scheme;
    uint32_t k0, k1;
    uint64_t t;
    uint32_t M;
    double delta, u0d;
    const double *u; // multinomial: the sorted uniforms of the step

    __device__ __forceinline__ double operator()(uint64_t j) const
    {
        if (scheme == MCKL_HIP_SYSTEMATIC)
            return __dadd_rn(__dmul_rn((double)j, delta), u0d);
        return __dmul_rn(delta, __dadd_rn(smc_uniform(k0, k1, t, j), (double)j));
    }

    // #{j < M : u_j < S}, from the guess floor(S M) by walking the
    // non-decreasing sequence.  For stratified uniforms
    // fl(delta j) <= u_j <= fl(delta (j+1)) (rounding is monotone and the
    // stratum variate lies in [0, 1)), so a step whose answer those bounds
    // already decide skips generating u_j: typically one evaluation per
    // count instead of two or three.  The count is the same either way.
    __device__ __forceinline__ uint32_t count_less(double S) const
    {
        if (scheme == MCKL_HIP_MULTINOMIAL) {
            // first j with u_j >= S (binary search: monotone in S)
            uint32_t lo = 0, hi = M;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo) / 2;
                if (u[mid] < S)
                    lo = mid + 1;
                else
                    hi = mid;
            }
            return lo;
        }
        double g = floor(__dmul_rn(S, (double)M));
        uint32_t c = g <= 0.0 ? 0u : (g >= (double)M ? M : (uint32_t)g);
        const bool strat = scheme == MCKL_HIP_STRATIFIED;
        while (c > 0) {
            if (strat && __dmul_rn(delta, (double)c) < S)
                break; // u_{c-1} <= fl(delta c) < S
            if ((*this)(c - 1) < S)
                break;
            --c;
        }
        while (c < M) {
            if (strat && __dmul_rn(delta, (double)c) >= S)
                break; // u_c >= fl(delta c) >= S
            if ((*this)(c) >= S)
                break;
            ++c;
        }
        return c;
    }

    // The same count without the walk in all but rounding-edge cases.  The
    // uniforms satisfy fl(delta j) <= u_j <= fl(delta (j+1)) (systematic:
    // u0d < delta), so in exact arithmetic the count is the guess
    // c = floor(S M) or c + 1; it is c + [u_c < S] whenever u_{c-1} < S and
    // u_{c+1} >= S, which the bounds above often decide without evaluating
    // u.  When either check fails (S M within rounding of an integer) the
    // walk gives the count.  Same result as count_less, always.
    __device__ __forceinline__ uint32_t count_fast(double S) const
    {
        if (scheme == MCKL_HIP_MULTINOMIAL)
            return count_less(S);
        const double g = floor(__dmul_rn(S, (double)M));
        const uint32_t c = g <= 0.0 ? 0u : (g >= (double)M ? M : (uint32_t)g);
        const bool strat = scheme == MCKL_HIP_STRATIFIED;
        const double blo = __dmul_rn(delta, (double)c);       // <= u_c
        const double bhi = __dmul_rn(delta, (double)c + 1.0); // u_c <= bhi
        const bool lo_ok = c == 0 || (strat && blo < S) || (*this)(c - 1) < S;
        const bool hi_ok = c + 1 >= M || (strat && bhi >= S) ||
            (*this)(c + 1) >= S;                            // u_{c+1} >= bhi
        uint32_t in = 0;
        if (c < M) {
            if (strat && bhi < S)
                in = 1;
            else if (!(strat && blo >= S))
                in = (*this)(c) < S ? 1u : 0u;
        }
        if (lo_ok && hi_ok)
            return c + in;
        return count_less(S);
    }

    // The count for every S within mg of Sf when they all agree (see
    // count_sys_guarded / count_strat_guarded); false when a uniform may lie
    // within mg of Sf, and always for the multinomial scheme.
    __device__ __forceinline__ bool count_guarded(
        double Sf, double mg, uint32_t &c) const
    {
        if (scheme == MCKL_HIP_SYSTEMATIC)
            return count_sys_guarded(Sf, mg, u0d, M, c);
        if (scheme != MCKL_HIP_STRATIFIED)
            return false;
        const int g = count_strat_guarded(Sf, mg, M, c);
        if (g != 2)
            return g == 1;
        const double uc = (*this)(c);
        if (uc < __dsub_rn(Sf, mg)) {
            ++c;
            return true;
        }
        return uc >= __dadd_rn(Sf, mg);
    }
};

// relative bound on |S - Sf| for the double evaluation of a cumulative
// weight in tile_counts_c: 64 unit roundoffs (the evaluation's own bound is
// about 17)
constexpr double kCountMargin = 64.0 * kUnitRoundoff;

// The tile-local part of tile_counts_c, which needs no value from other
// tiles: a_i = exp_f(lw_i - m_k) of the thread's particles (tile_a; 0 past
// N) and the thread's exclusive double-double prefix of a within the tile
// (toff; tile_sums_l returns the same prefix).
__device__ __forceinline__ void tile_a(const SmcArgs &A, uint32_t tile,
    const float (&l)[kItems], float mk, int gt, float (&a)[kItems])
{
    const uint64_t i0 = tile_first(tile, gt);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const float e = smc_exp_f(__fsub_rn(l[k], mk));
        a[k] = i0 + k < A.N ? e : 0.0f;
    }
}

__device__ __forceinline__ void tile_counts_local(const SmcArgs &A,
    uint32_t tile, const float (&l)[kItems], float mk, int gt, double *ldsd,
    float (&a)[kItems], dd &toff)
{
    tile_a(A, tile, l, mk, gt, a);
    dd loc{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        loc = dd_add(loc, dd{(double)a[k], 0.0});
    dd total;
    toff = group_excl_scan_dd(loc, gt, ldsd, total);
}

// a thread's exclusive prefixes of the tile's extra-copy and zero-slot
// counts and the tile's totals (tile_counts_c can hand them to
// tile_index_local, which would otherwise scan them again)
struct TileScan {
    uint32_t ex, zx, te, tz;
};

// Cumulative normalised weights S_j = round(prefix_j(e)) * (1/S), prefix in
// double-double: the tile's offset (tile_off, from the exclusive scan of the
// scaled tile totals) plus r_k times the in-tile prefix of a_i; the last
// particle of a tile takes the next tile's offset (next_hi) so that
// neighbouring tiles agree on boundary values, and the last particle overall
// takes cnt = M (resample.hpp:405).  The counts cnt_j = #{u < S_j} are then
// made non-decreasing and confined to [count(tile_off), count(next_hi)], the
// values the neighbouring tiles use at the shared boundaries, so the
// replication numbers are never negative and sum to M whatever the rounding
// of the prefix sums (the clamp changes nothing when those are monotone).
// Returns the thread's counts in c, the count before its first particle in
// prev, and the tile sums of the extra-copy and zero-slot counts through
// esum/zsum.  (a, toff: from tile_counts_local)
__device__ __forceinline__ void tile_counts_c(const SmcArgs &A, uint32_t tile,
    const float (&a)[kItems], const dd toff, double rk, double invS,
    dd tile_off, double next_hi, const UniformSeq &U, int gt,
    uint32_t *ldsu, uint32_t (&c)[kItems], uint32_t &prev_out,
    uint32_t &esum, uint32_t &zsum, TileScan *scan = nullptr)
{
    const uint64_t i0 = tile_first(tile, gt);
    const uint64_t tend = (uint64_t)(tile + 1) * kTile < A.N
        ? (uint64_t)(tile + 1) * kTile
        : A.N;
    // boundary counts shared with the neighbouring tiles (lane 0: the
    // tile's first boundary, lane 1: its last; one copy of the count code)
    const uint32_t cb = U.count_fast(
        __dmul_rn((lane_id() & 1) ? next_hi : tile_off.hi, invS));
    const uint32_t clo = tile == 0
        ? 0u
        : (uint32_t)__builtin_amdgcn_readlane((int)cb, 0);
    const uint32_t chi = tend == A.N
        ? U.M
        : (uint32_t)__builtin_amdgcn_readlane((int)cb, 1);
    // P_i = tile_off + r_k (toff + sum_{j <= i} a_j): in double first
    // (relative error below 17 unit roundoffs, all terms non-negative); the
    // double-double value rounded once only where a uniform may lie within
    // the bound of the double one (count_guarded), so the count is that of
    // the rounded double-double prefix in every case.  Those (rare) items
    // are marked in `redo` and counted afterwards in one rolled loop, so the
    // exact path's code is not repeated per item.
    double rd = toff.hi;
    uint32_t redo = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        rd = __dadd_rn(rd, (double)a[k]);
        uint32_t v = chi;
        if (i < tend - 1) {
            const double Pf = __dadd_rn(__dmul_rn(rd, rk), tile_off.hi);
            const double Sf = __dmul_rn(Pf, invS);
            if (!U.count_guarded(Sf, __dmul_rn(Sf, kCountMargin), v))
                redo |= 1u << k;
        }
        c[k] = v;
    }
    if (redo) {
        dd run = toff;
#pragma unroll 1
        for (int k = 0; k < kItems; ++k) {
            float ak = a[0];
#pragma unroll
            for (int q = 1; q < kItems; ++q)
                ak = k == q ? a[q] : ak;
            run = dd_add(run, dd{(double)ak, 0.0});
            if ((redo >> k) & 1u) {
                const double P = dd_add(tile_off, dd_mul_d(run, rk)).hi;
                const uint32_t v = U.count_fast(__dmul_rn(P, invS));
#pragma unroll
                for (int q = 0; q < kItems; ++q)
                    c[q] = k == q ? v : c[q];
            }
        }
    }
    uint32_t cmax = clo;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < A.N) {
            const uint32_t v = clamp_count(c[k], clo, chi);
            cmax = v > cmax ? v : cmax;
            c[k] = cmax;
        } else {
            c[k] = 0;
        }
    }
    // counts of the threads before this one: make the sequence monotone
    const uint32_t before = group_excl_max_u(cmax, clo, gt, ldsu);
    uint32_t prev = before;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (i0 + k < A.N && c[k] < before)
            c[k] = before;
    prev_out = before;
    uint32_t es = 0, zs = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < A.N) {
            const uint32_t rep = rep_from_counts(c[k], prev);
            es += extra_copies(rep);
            zs += rep == 0 ? 1 : 0;
        }
        prev = c[k];
    }
    if (scan) { // uniform over the group
        scan->ex = group_excl_scan_u(es, gt, ldsu, scan->te);
        scan->zx = group_excl_scan_u(zs, gt, ldsu + 4, scan->tz);
        esum = scan->te;
        zsum = scan->tz;
        return;
    }
    group_sum2_u(es, zs, gt, ldsu);
    esum = es;
    zsum = zs;
}

// tile_counts_c, the counts stored to cnt
__device__ __forceinline__ void tile_counts_l(const SmcArgs &A, uint32_t tile,
    const float (&l)[kItems], float mk, double rk, double invS, dd tile_off,
    double next_hi, const UniformSeq &U, uint32_t *cnt, int gt, double *ldsd,
    uint32_t *ldsu, uint32_t &esum, uint32_t &zsum)
{
    uint32_t c[kItems], prev;
    float a[kItems];
    dd toff;
    tile_counts_local(A, tile, l, mk, gt, ldsd, a, toff);
    tile_counts_c(A, tile, a, toff, rk, invS, tile_off, next_hi, U, gt,
        ldsu, c, prev, esum, zsum);
    store8(cnt, tile_first(tile, gt), A.N, c);
}

__device__ __forceinline__ void tile_counts(const SmcArgs &A, uint32_t tile,
    const float *lw, float mk, double rk, double invS, dd tile_off,
    double next_hi, const UniformSeq &U, uint32_t *cnt, int gt, double *ldsd,
    uint32_t *ldsu, uint32_t &esum, uint32_t &zsum)
{
    float l[kItems];
    load8(lw, tile_first(tile, gt), A.N, l, 0.0f);
    tile_counts_l(A, tile, l, mk, rk, invS, tile_off, next_hi, U, cnt, gt,
        ldsd, ldsu, esum, zsum);
}

// Residual resampling in the samplers (ResampleAlgorithm<U01Seq, true>,
// resample.hpp:333-364 and 515-593).  With the normalised weight
// W_i = fl(e_i * (1/S)), M W_i = I_i + r_i (modf: I_i integral, r_i in
// [0, 1), both exact); R = M - sum I.  Particle i keeps I_i copies and the
// remaining R are drawn with the base scheme's R uniforms against the
// cumulative residuals normalised by their total (double-double prefix
// rounded once, as for the plain schemes).
__device__ __forceinline__ void resid_part(
    double e, double invS, double M, uint32_t &I, double &r)
{
    const double w = __dmul_rn(M, __dmul_rn(e, invS));
    const double fw = floor(w);
    I = (uint32_t)fw;
    r = __dsub_rn(w, fw);
}

// the samplers' weights e_i = a_i r_k of a thread's particles (0 past N)
__device__ __forceinline__ void tile_weights(const SmcArgs &A, uint32_t tile,
    const float (&l)[kItems], float mk, double rk, int gt,
    double (&e)[kItems])
{
    const uint64_t i0 = tile_first(tile, gt);
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        e[k] = i0 + k < A.N
            ? __dmul_rn((double)smc_exp_f(__fsub_rn(l[k], mk)), rk)
            : 0.0;
}

// the tile's sum of I_i and double-double total of r_i, from the thread's
// weights e (the FP32 samplers' tile_weights, the gradient's FP64 weights)
__device__ __forceinline__ void tile_residual_sums_e(const SmcArgs &A,
    uint32_t tile, const double (&ew)[kItems], double invS, int gt,
    double *ldsd, uint32_t *ldsu, uint32_t &isum, dd &rtot)
{
    const uint64_t i0 = tile_first(tile, gt);
    const double M = (double)A.N;
    uint32_t is = 0;
    dd loc{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < A.N) {
            const double e = ew[k];
            uint32_t I;
            double r;
            resid_part(e, invS, M, I, r);
            is += I;
            loc = dd_add(loc, dd{r, 0.0});
        }
    }
    uint32_t zero = 0;
    group_sum2_u(is, zero, gt, ldsu);
    isum = is;
    (void)group_excl_scan_dd(loc, gt, ldsd, rtot);
}

__device__ __forceinline__ void tile_residual_sums(const SmcArgs &A,
    uint32_t tile, const float (&l)[kItems], float mk, double rk, double invS,
    int gt, double *ldsd, uint32_t *ldsu, uint32_t &isum, dd &rtot)
{
    double e[kItems];
    tile_weights(A, tile, l, mk, rk, gt, e);
    tile_residual_sums_e(A, tile, e, invS, gt, ldsd, ldsu, isum, rtot);
}

// Cumulative counts of a residual step: cnt_i = (sum_{j <= i} I_j) +
// (residual count of i), the residual counts #{u_j < P^r_i (1/Rtot)} over
// the R uniforms of U (U.M = R), with P^r the double-double prefix of the
// residuals (roff: the tile's offset, rnext_hi: the next tile's) and the
// same boundary clamp as tile_counts_l.  Ioff: sum of I over the tiles
// before this one.  Returns the tile's extra-copy and zero-slot counts.
__device__ __forceinline__ void tile_rcounts_e(const SmcArgs &A,
    uint32_t tile, const double (&ew)[kItems], double invS, uint32_t Ioff,
    dd roff, double rnext_hi, double invR, const UniformSeq &U, uint32_t *cnt,
    int gt, double *ldsd, uint32_t *ldsu, uint32_t &esum, uint32_t &zsum)
{
    const uint64_t i0 = tile_first(tile, gt);
    const uint64_t tend = (uint64_t)(tile + 1) * kTile < A.N
        ? (uint64_t)(tile + 1) * kTile
        : A.N;
    const double M = (double)A.N;
    const uint32_t R = U.M;
    const uint32_t clo = (tile == 0 || R == 0)
        ? 0u
        : U.count_fast(__dmul_rn(roff.hi, invR));
    const uint32_t chi = tend == A.N
        ? R
        : (R == 0 ? 0u : U.count_fast(__dmul_rn(rnext_hi, invR)));
    uint32_t I[kItems];
    double r[kItems];
    uint32_t is = 0;
    dd loc{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        I[k] = 0;
        r[k] = 0.0;
        if (i0 + k < A.N)
            resid_part(ew[k], invS, M, I[k], r[k]);
        is += I[k];
        loc = dd_add(loc, dd{r[k], 0.0});
    }
    dd total;
    const dd toff = group_excl_scan_dd(loc, gt, ldsd, total);
    uint32_t itot;
    const uint32_t iex = group_excl_scan_u(is, gt, ldsu, itot);
    uint32_t cr[kItems];
    dd run = toff;
    uint32_t cmax = clo;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        run = dd_add(run, dd{r[k], 0.0});
        cr[k] = 0;
        if (i < A.N) {
            uint32_t v;
            if (i == tend - 1)
                v = chi;
            else if (R == 0)
                v = 0;
            else
                v = U.count_fast(__dmul_rn(dd_add(roff, run).hi, invR));
            v = clamp_count(v, clo, chi);
            cmax = v > cmax ? v : cmax;
            cr[k] = cmax;
        }
    }
    const uint32_t before = group_excl_max_u(cmax, clo, gt, ldsu);
    uint32_t ic = Ioff + iex;
    uint32_t prev = ic + before; // count of the particle before i0
    uint32_t c[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        ic += I[k];
        const uint32_t q = cr[k] < before ? before : cr[k];
        c[k] = ic + q;
    }
    store8(cnt, i0, A.N, c);
    uint32_t es = 0, zs = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        if (i0 + k < A.N) {
            const uint32_t rep = rep_from_counts(c[k], prev);
            es += extra_copies(rep);
            zs += rep == 0 ? 1 : 0;
        }
        prev = c[k];
    }
    group_sum2_u(es, zs, gt, ldsu);
    esum = es;
    zsum = zs;
}

__device__ __forceinline__ void tile_rcounts(const SmcArgs &A, uint32_t tile,
    const float (&l)[kItems], float mk, double rk, double invS, uint32_t Ioff,
    dd roff, double rnext_hi, double invR, const UniformSeq &U, uint32_t *cnt,
    int gt, double *ldsd, uint32_t *ldsu, uint32_t &esum, uint32_t &zsum)
{
    double e[kItems];
    tile_weights(A, tile, l, mk, rk, gt, e);
    tile_rcounts_e(A, tile, e, invS, Ioff, roff, rnext_hi, invR, U, cnt, gt,
        ldsd, ldsu, esum, zsum);
}

// Stable rep->index (resample.hpp:416-466, M = N): idx_i = i when rep_i > 0;
// zero slots take extra copies in ascending source order.  Writes the
// inclusive extra-copy prefix E_i, idx_i = i for every slot, and
// zpos[r] = i for the zero slot of rank r; tile_scatter_ranks then overwrites the
// zero slots with their sources.
// the thread's 8 counts and the count before them
__device__ __forceinline__ void load_counts(const SmcArgs &A, uint32_t tile,
    const uint32_t *cnt, int gt, uint32_t (&c)[kItems], uint32_t &prev)
{
    const uint64_t i0 = tile_first(tile, gt);
    load8(cnt, i0, A.N, c, 0u);
    prev = (i0 == 0 || i0 > A.N) ? 0u : cnt[i0 - 1];
}

__device__ __forceinline__ void tile_index_c(const SmcArgs &A, uint32_t tile,
    const uint32_t (&c)[kItems], uint32_t prev, uint32_t eoff, uint32_t zoff,
    uint32_t *Eincl, uint32_t *idx, uint32_t *zpos, int gt, uint32_t *ldsu)
{
    const uint64_t i0 = tile_first(tile, gt);
    uint32_t rep[kItems];
    uint32_t es = 0, zs = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        rep[k] = 0;
        if (i < A.N) {
            rep[k] = rep_from_counts(c[k], prev);
            prev = c[k];
            es += extra_copies(rep[k]);
            zs += rep[k] == 0 ? 1 : 0;
        }
    }
    uint32_t te, tz;
    uint32_t ex = group_excl_scan_u(es, gt, ldsu, te) + eoff;
    uint32_t zx = group_excl_scan_u(zs, gt, ldsu + 4, tz) + zoff;
    uint32_t E[kItems], own[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        ex += extra_copies(rep[k]);
        E[k] = ex;
        own[k] = (uint32_t)i; // zero slots: overwritten by tile_scatter_ranks
        if (i < A.N && rep[k] == 0) {
            zpos[zx] = (uint32_t)i;
            ++zx;
        }
    }
    store8(Eincl, i0, A.N, E);
    store8(idx, i0, A.N, own);
}

__device__ __forceinline__ void tile_index(const SmcArgs &A, uint32_t tile,
    const uint32_t *cnt, uint32_t eoff, uint32_t zoff, uint32_t *Eincl,
    uint32_t *idx, uint32_t *zpos, int gt, uint32_t *ldsu)
{
    uint32_t c[kItems], prev;
    load_counts(A, tile, cnt, gt, c, prev);
    tile_index_c(A, tile, c, prev, eoff, zoff, Eincl, idx, zpos, gt, ldsu);
}

// Stable rep->index with the zero slots' parents left as codes, for the
// plain sampler: idx_i = i when rep_i > 0, else kZeroFlag | (zoff + the
// slot's rank among the tile's zero slots); src[eoff + r] = the source of
// the tile's extra copy r, the first particle j of the tile with E_j > r (E
// the tile's inclusive extra-copy prefix, held in LDS).  The next move
// resolves a code through src (resolve_parent), so the zero slot of rank r
// takes the source of extra copy r: the order of resample.hpp:416-466.
// eoff/zoff: the extra-copy and zero-slot totals of the tiles before this
// one.  Two parts, so that a caller can wait for eoff/zoff between them:
// tile_index_local does everything the offsets do not enter (the codes'
// local ranks, E in LDS, and, when the tile holds at most kTile extra
// copies, each rank's source in LDS), tile_index_store adds the offsets
// and writes idx and src.  Each contains workgroup barriers: every thread
// of the workgroup calls them.
struct TileIndex {
    uint32_t code[kItems]; // own index, or kZeroFlag | local zero rank
    uint32_t n;            // the tile's particles
    uint32_t end;          // the tile's extra copies
};

// four independent rank -> source searches at once (their LDS round trips
// overlap); ranks at or past `end` are skipped
__device__ __forceinline__ void rank_sources4(const uint32_t *Es, uint32_t n,
    uint32_t end, uint32_t r0, uint32_t stride, uint32_t (&j)[4])
{
    uint32_t lo[4], hi[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        lo[q] = 0;
        hi[q] = r0 + q * stride < end ? n - 1 : 0u;
    }
    // at most ceil(log2(kTile)) halvings
#pragma unroll 1
    for (int it = 0; it < 12; ++it) {
        bool more = false;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (lo[q] < hi[q]) {
                const uint32_t mid = (lo[q] + hi[q]) >> 1;
                if (Es[mid] > r0 + q * stride)
                    hi[q] = mid;
                else
                    lo[q] = mid + 1;
                more = true;
            }
        }
        if (!more)
            break;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        j[q] = lo[q];
}

__device__ __forceinline__ void tile_index_local(const SmcArgs &A,
    uint32_t tile, const uint32_t (&c)[kItems], uint32_t prev, int gt,
    uint32_t *ldsu, uint32_t *Es, uint32_t *srcl, TileIndex &ti,
    const TileScan *scan = nullptr)
{
    const uint64_t t0 = (uint64_t)tile * kTile;
    const uint64_t i0 = tile_first(tile, gt);
    uint32_t rep[kItems];
    uint32_t es = 0, zs = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        rep[k] = 0;
        if (i0 + k < A.N) {
            rep[k] = rep_from_counts(c[k], prev);
            prev = c[k];
            es += extra_copies(rep[k]);
            zs += rep[k] == 0 ? 1 : 0;
        }
    }
    uint32_t te, tz, ex, zx;
    if (scan) { // the counting stage's scans of the same counts
        ex = scan->ex;
        zx = scan->zx;
        te = scan->te;
        tz = scan->tz;
    } else {
        ex = group_excl_scan_u(es, gt, ldsu, te);
        zx = group_excl_scan_u(zs, gt, ldsu + 4, tz);
    }
    ti.n = t0 >= A.N ? 0u
        : (uint32_t)(A.N - t0 < (uint64_t)kTile ? A.N - t0 : kTile);
    ti.end = te;
    const bool small = te <= (uint32_t)kTile; // uniform over the group
    if (small) {
        reinterpret_cast<uint4 *>(srcl)[2 * gt] = make_uint4(0, 0, 0, 0);
        reinterpret_cast<uint4 *>(srcl)[2 * gt + 1] = make_uint4(0, 0, 0, 0);
        __syncthreads();
    }
    uint32_t E[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        // a source's extra copies are the ranks [ex, ex + rep - 1): it
        // marks the first with its own index
        if (small && rep[k] > 1 && ex < (uint32_t)kTile)
            srcl[ex] = (uint32_t)(gt * kItems + k);
        ex += extra_copies(rep[k]);
        E[k] = ex;
        ti.code[k] = (uint32_t)i;
        if (i < A.N && rep[k] == 0)
            ti.code[k] = kZeroFlag | zx++;
    }
    reinterpret_cast<uint4 *>(Es)[2 * gt] = make_uint4(E[0], E[1], E[2], E[3]);
    reinterpret_cast<uint4 *>(Es)[2 * gt + 1] =
        make_uint4(E[4], E[5], E[6], E[7]);
    __syncthreads();
    if (small) {
        // every rank's source: the last mark at or before it (sources mark
        // increasing ranks in particle order), by a max-scan over the ranks
        const uint4 a = reinterpret_cast<const uint4 *>(srcl)[2 * gt];
        const uint4 b = reinterpret_cast<const uint4 *>(srcl)[2 * gt + 1];
        uint32_t v[kItems] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int k = 1; k < kItems; ++k)
            v[k] = v[k] > v[k - 1] ? v[k] : v[k - 1];
        const uint32_t before = group_excl_max_u(v[kItems - 1], 0u, gt, ldsu);
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            v[k] = v[k] > before ? v[k] : before;
        reinterpret_cast<uint4 *>(srcl)[2 * gt] =
            make_uint4(v[0], v[1], v[2], v[3]);
        reinterpret_cast<uint4 *>(srcl)[2 * gt + 1] =
            make_uint4(v[4], v[5], v[6], v[7]);
        __syncthreads();
    }
}

__device__ __forceinline__ void tile_index_store(const SmcArgs &A,
    uint32_t tile, const TileIndex &ti, uint32_t eoff, uint32_t zoff,
    uint32_t *idx, uint32_t *src, int gt, const uint32_t *Es,
    const uint32_t *srcl)
{
    const uint64_t t0 = (uint64_t)tile * kTile;
    const uint64_t i0 = tile_first(tile, gt);
    uint32_t code[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        code[k] = (ti.code[k] & kZeroFlag) ? ti.code[k] + zoff : ti.code[k];
    store8_nt(idx, i0, A.N, code);
    // ranks with a slot in src: below N - eoff (the loop stays bounded even
    // for counts that are not a replication vector)
    const uint32_t end = rank_end(ti.end, eoff < A.N ? A.N - eoff : 0u);
    if (ti.end <= (uint32_t)kTile) { // sources from tile_index_local
        for (uint32_t r = gt; r < end; r += kGroup)
            __builtin_nontemporal_store(
                (uint32_t)(t0 + srcl[r]), src + eoff + r);
    } else {
        for (uint32_t r0 = gt; r0 < end; r0 += 4 * kGroup) {
            uint32_t j[4];
            rank_sources4(Es, ti.n, end, r0, kGroup, j);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (r0 + q * kGroup < end)
                    src[eoff + r0 + q * kGroup] = (uint32_t)(t0 + j[q]);
        }
    }
}

// Zero slots take their sources.  Source j with rep_j > 1 owns the extra
// copies of ranks [E_j - (rep_j - 1), E_j), so rank r of the tile's range
// [E_{t0-1}, E_{t0+n-1}) belongs to the first particle j of the tile with
// E_j > r, and idx[zpos[r]] = j.  The ranks go round-robin to the group's
// threads and each finds its source by binary search in the tile's E held in
// LDS (Es, kTile words), so the tile's work is spread evenly over its
// threads however the copies are distributed among its particles, and the
// zpos loads of consecutive threads are contiguous.  A tile at or past the
// end (the batched forms' padding rounds) writes nothing.  Contains two
// workgroup barriers: every thread of the workgroup calls it.
__device__ __forceinline__ void tile_scatter_ranks(const SmcArgs &A,
    uint32_t tile, const uint32_t *__restrict__ Eincl,
    const uint32_t *__restrict__ zpos, uint32_t *__restrict__ idx, int gt,
    uint32_t *Es)
{
    const uint64_t t0 = (uint64_t)tile * kTile;
    const bool live = t0 < A.N;
    const uint32_t n = !live ? 0u
        : (uint32_t)(A.N - t0 < (uint64_t)kTile ? A.N - t0 : kTile);
    const uint64_t i0 = tile_first(tile, gt);
    uint32_t E[kItems];
    load8(Eincl, i0, A.N, E, 0u);
    const uint32_t beg = (!live || t0 == 0) ? 0u : Eincl[t0 - 1];
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        Es[gt * kItems + k] = E[k];
    __syncthreads();
    // ranks and slots stay inside the arrays even if the counts were not a
    // replication vector (they are: the counts are clamped monotone)
    const uint32_t end = live ? rank_end(Es[n - 1], A.N) : 0u;
    for (uint32_t r = beg + gt; r < end; r += kGroup) {
        const uint32_t lo = rank_source(Es, n, r); // Es[n - 1] >= end > r
        const uint32_t z = zpos[r];
        if (z < A.N)
            idx[z] = (uint32_t)(t0 + lo);
    }
    __syncthreads();
}

// Multinomial uniforms of step t (u01_rand_sorted, resample.hpp:63-81,
// 216-235, on the step's uniform stream U_j = smc_uniform(t, j)):
// lm_j = sum_{i <= j} log(U_i) / (N - i), u_j = 1 - exp(lm_j), ascending.
// The sum is a double-double prefix over tiles (tile totals here, offsets by
// tiles_offset_dd/tiles_scan_dd), rounded once per element.
// M: the number of uniforms (N, or R of a residual step)
__device__ __forceinline__ void tile_mterms(const SmcArgs &A, uint32_t tile,
    uint64_t t, int gt, double *ldsd, dd &ttot, uint32_t M)
{
    const uint64_t i0 = tile_first(tile, gt);
    dd loc{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t j = i0 + k;
        if (j < M) {
            const double a = __ddiv_rn(
                log(smc_uniform(A.k0, A.k1, t, j)), (double)(M - j));
            loc = dd_add(loc, dd{a, 0.0});
        }
    }
    (void)group_excl_scan_dd(loc, gt, ldsd, ttot);
}

__device__ __forceinline__ void tile_mterms(const SmcArgs &A, uint32_t tile,
    uint64_t t, int gt, double *ldsd, dd &ttot)
{
    tile_mterms(A, tile, t, gt, ldsd, ttot, A.N);
}

__device__ __forceinline__ void tile_muniforms(const SmcArgs &A,
    uint32_t tile, uint64_t t, dd tile_off, double *u, int gt, double *ldsd,
    uint32_t M)
{
    const uint64_t i0 = tile_first(tile, gt);
    double a[kItems];
    dd loc{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t j = i0 + k;
        a[k] = j < M ? __ddiv_rn(log(smc_uniform(A.k0, A.k1, t, j)),
                           (double)(M - j))
                     : 0.0;
        loc = dd_add(loc, dd{a[k], 0.0});
    }
    dd tot;
    const dd base = dd_add(tile_off, group_excl_scan_dd(loc, gt, ldsd, tot));
    dd run{0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t j = i0 + k;
        run = dd_add(run, dd{a[k], 0.0});
        if (j < M)
            u[j] = __dsub_rn(1.0, exp(dd_add(base, run).hi));
    }
}

__device__ __forceinline__ void tile_muniforms(const SmcArgs &A,
    uint32_t tile, uint64_t t, dd tile_off, double *u, int gt, double *ldsd)
{
    tile_muniforms(A, tile, t, tile_off, u, gt, ldsd, A.N);
}

// finalize one step from the reduced sums; returns the resampling flag
struct StepOut {
    double logz_inc, ess;
    int resample;
};

__device__ __forceinline__ StepOut step_finalize(const SmcArgs &A, float m,
    double S, double S2, double C)
{
    StepOut o;
    o.logz_inc = __dsub_rn(__dadd_rn((double)m, log(S)), log(C));
    o.ess = __ddiv_rn(__dmul_rn(S, S), S2);
    o.resample = o.ess < A.tauN ? 1 : 0;
    return o;
}

__device__ __forceinline__ UniformSeq make_uniform_seq(
    const SmcArgs &A, uint64_t t, const double *u = nullptr)
{
    UniformSeq U;
    U.u = u;
    U.scheme = A.scheme;
    U.k0 = A.k0;
    U.k1 = A.k1;
    U.t = t;
    U.M = A.N;
    U.delta = __ddiv_rn(1.0, (double)A.N);
    U.u0d = A.scheme == MCKL_HIP_SYSTEMATIC ? __dmul_rn(smc_uniform(A.k0, A.k1, t, 0), U.delta)
                          : 0.0;
    return U;
}

} // namespace mckl_hip
