// smc_grad.hpp -- batched gradient of the SMC log-likelihood with a
// hand-written adjoint (no reference counterpart: the reference exposes no
// log-likelihood, core/weight.hpp:198-227, and no derivatives).
//
// One workgroup per parameter point, FP64 particle states and log-weights.
// The forward pass follows the samplers of smc.hip step for step (same noise
// streams, same resampling decisions and stable rep->index transform) and
// stores a checkpoint every K steps.  The backward pass walks the segments in
// reverse: it recomputes a segment from its checkpoint (the kernels are
// deterministic, so ancestors come back identical without being stored) and
// sweeps the adjoint through it.
//
// The differentiated function, with ancestors and resampling decisions held
// at their values at theta0 (the point of the forward pass):
//   F(theta) = sum_t [ LSE(lw_t) - LSE(carry_t) ]
//   lw_t,i   = carry_t,i + log g(y_t | x_t,i)
//   x_t,i    = a x_{t-1,p(i)} + s eps_t,i,  x_0,i = sd0 eps_0,i
//   carry    = lw_{t-1,i} - m_{t-1}                  (no resampling at t-1)
//            = LWn_{t-1,p(i)}(theta) - LWn_{t-1,p(i)}(theta0)  (resampled)
// with LWn the normalised log-weights.  F(theta0) is the estimate log Z-hat;
// the resampled weights carry their dependence on theta (value 1, gradient
// d log W), which makes grad F a consistent estimate of the score.  The
// replay entry point evaluates F at any theta with a tape recorded at
// theta0, for finite-difference checks.
#pragma once

#include "common.hpp"
#include "smc_tiles.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace mckl_hip {

constexpr int kMaxTilesG = 64; // N <= 131072 per point

struct GradArgs {
    int model;                  // 1 LG (a, sx, sy), 2 SV (phi, sigma, beta)
    double a, s, c;             // theta
    double sd0, dsd0_a, dsd0_s; // initial scale and its derivatives
    double cst;                 // -log c - log(2 pi)/2
    uint32_t k0, k1;            // particle RNG key
    uint32_t N, ntiles, T, K;
    double tauN;
    int scheme;
    const double *y;
};

// Per-slot layout: one slot holds the state after step t (x, lw, and the
// resampling structures that give the parents of step t+1).
struct SlotPtr {
    double *x, *lw;
    uint32_t *par, *E, *zpos; // resolved parents of step t+1's slots
    double *scal; // m, S, lse, R
};

__device__ __forceinline__ SlotPtr slot_at(double *base, size_t slot_bytes,
    uint32_t k, uint32_t Np)
{
    char *p = reinterpret_cast<char *>(base) + (size_t)k * slot_bytes;
    SlotPtr s;
    s.x = reinterpret_cast<double *>(p);
    s.lw = s.x + Np;
    s.par = reinterpret_cast<uint32_t *>(s.lw + Np);
    s.E = s.par + Np;
    s.zpos = s.E + Np;
    s.scal = reinterpret_cast<double *>(s.zpos + Np);
    return s;
}

__device__ __forceinline__ void gmodel(const GradArgs &G, double y, double x,
    double &g, double &gx, double &gc)
{
    if (G.model == 1) {
        const double ic = 1.0 / G.c;
        const double z = (y - x) * ic;
        g = -0.5 * z * z + G.cst;
        gx = z * ic;
        gc = (z * z - 1.0) * ic;
    } else {
        const double q = y * y * exp(-x) / (G.c * G.c);
        g = -0.5 * q - 0.5 * x + G.cst;
        gx = 0.5 * q - 0.5;
        gc = (q - 1.0) / G.c;
    }
}

__device__ __forceinline__ double eps_at(const GradArgs &G, uint64_t i,
    uint64_t t)
{
    float e0, e1;
    smc_normal_pair(G.k0, G.k1, i >> 1, t, e0, e1);
    return (double)((i & 1) ? e1 : e0);
}

// The 8 doubles of a thread's particles: four 16-byte loads when all are
// below N (slot arrays are whole tiles, 16-byte aligned), else element loads
// with `fill` past N.  Issuing a thread's loads together (then computing)
// keeps them in flight at once instead of one round trip per particle.
__device__ __forceinline__ void load8d(const double *p, uint64_t i0,
    uint64_t N, double (&v)[kItems], double fill)
{
    if (i0 + kItems <= N) {
        const double2 *q = reinterpret_cast<const double2 *>(p + i0);
#pragma unroll
        for (int j = 0; j < kItems / 2; ++j) {
            const double2 a = q[j];
            v[2 * j] = a.x;
            v[2 * j + 1] = a.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            v[k] = i0 + k < N ? p[i0 + k] : fill;
    }
}

__device__ __forceinline__ void store8d(double *p, uint64_t i0, uint64_t N,
    const double (&v)[kItems])
{
    if (i0 + kItems <= N) {
        double2 *q = reinterpret_cast<double2 *>(p + i0);
#pragma unroll
        for (int j = 0; j < kItems / 2; ++j)
            q[j] = make_double2(v[2 * j], v[2 * j + 1]);
    } else {
#pragma unroll
        for (int k = 0; k < kItems; ++k)
            if (i0 + k < N)
                p[i0 + k] = v[k];
    }
}

// ---- forward pieces (FP64 counterparts of smc_tiles.hpp) ----
__device__ __forceinline__ double gtile_move(const GradArgs &G, uint32_t t,
    uint32_t tile, const SlotPtr &prev, const SlotPtr &cur, int Rprev,
    double mprev, const double *lwn_ref, const uint32_t *tpar, int gt,
    double *ldsd)
{
    const uint64_t i0 = tile_first(tile, gt);
    const double y = G.y[t - 1];
    double lmax = -__builtin_huge_val();
    if (i0 + kItems <= G.N) {
        // a whole row of 8: the loads (and the parents' gathers) issued
        // together, the normals drawn while they are in flight; the same
        // arithmetic as the element loop below
        double xp[kItems], carry[kItems];
        float e[kItems];
        if (Rprev) {
            uint32_t p[kItems];
            load8(tpar ? tpar : prev.par, i0, G.N, p, 0u);
#pragma unroll
            for (int k = 0; k < kItems; k += 2)
                smc_normal_pair(G.k0, G.k1, (i0 + k) >> 1, t, e[k], e[k + 1]);
#pragma unroll
            for (int k = 0; k < kItems; ++k)
                xp[k] = prev.x[p[k]];
            if (lwn_ref) {
                const double l2 = prev.scal[2];
                double a[kItems], b[kItems];
#pragma unroll
                for (int k = 0; k < kItems; ++k) {
                    a[k] = prev.lw[p[k]];
                    b[k] = lwn_ref[p[k]];
                }
#pragma unroll
                for (int k = 0; k < kItems; ++k)
                    carry[k] = (a[k] - l2) - b[k];
            } else {
#pragma unroll
                for (int k = 0; k < kItems; ++k)
                    carry[k] = 0.0;
            }
        } else {
            double lv[kItems];
            load8d(prev.x, i0, G.N, xp, 0.0);
            load8d(prev.lw, i0, G.N, lv, 0.0);
#pragma unroll
            for (int k = 0; k < kItems; k += 2)
                smc_normal_pair(G.k0, G.k1, (i0 + k) >> 1, t, e[k], e[k + 1]);
#pragma unroll
            for (int k = 0; k < kItems; ++k)
                carry[k] = lv[k] - mprev;
        }
        double xo[kItems], lo[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const double x = G.a * xp[k] + G.s * (double)e[k];
            double g, gx, gc;
            gmodel(G, y, x, g, gx, gc);
            const double l = carry[k] + g;
            xo[k] = x;
            lo[k] = l;
            lmax = lmax < l ? l : lmax;
        }
        store8d(cur.x, i0, G.N, xo);
        store8d(cur.lw, i0, G.N, lo);
    } else
    // particles i0.. are consecutive from an even index: one normal pair per
    // two particles (eps_at's values, each pair drawn once)
    for (int k = 0; k < kItems; k += 2) {
        if (i0 + k >= G.N)
            break;
        float e[2];
        smc_normal_pair(G.k0, G.k1, (i0 + k) >> 1, t, e[0], e[1]);
        for (int q = 0; q < 2; ++q) {
            const uint64_t i = i0 + k + q;
            if (i >= G.N)
                break;
            double xp, carry;
            if (Rprev) {
                const uint32_t p = tpar ? tpar[i] : prev.par[i];
                xp = prev.x[p];
                // value of the carried weight: LWn(theta)[p] - LWn(theta0)[p]
                carry =
                    lwn_ref ? (prev.lw[p] - prev.scal[2]) - lwn_ref[p] : 0.0;
            } else {
                xp = prev.x[i];
                carry = prev.lw[i] - mprev;
            }
            const double x = G.a * xp + G.s * (double)e[q];
            double g, gx, gc;
            gmodel(G, y, x, g, gx, gc);
            const double l = carry + g;
            cur.x[i] = x;
            cur.lw[i] = l;
            lmax = lmax < l ? l : lmax;
        }
    }
    // group max of doubles
    lmax = wave_max_d(lmax);
    if ((gt & 63) == 0)
        ldsd[gt >> 6] = lmax;
    __syncthreads();
    double r = ldsd[0];
    for (int w = 1; w < kWaves; ++w)
        r = r < ldsd[w] ? ldsd[w] : r;
    __syncthreads();
    return r;
}

__device__ __forceinline__ void gtile_sums(const GradArgs &G, uint32_t tile,
    const double *lw, double m, int gt, double *ldsd, double &S, double &S2,
    dd &etot)
{
    const uint64_t i0 = tile_first(tile, gt);
    double s = 0.0, s2 = 0.0;
    dd loc{0.0, 0.0};
    double lv[kItems];
    load8d(lw, i0, G.N, lv, 0.0);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < G.N) {
            const double e = exp(lv[k] - m);
            s += e;
            s2 += e * e;
            loc = dd_add(loc, dd{e, 0.0});
        }
    }
    group_sum2(s, s2, gt, ldsd);
    S = s;
    S2 = s2;
    (void)group_excl_scan_dd(loc, gt, ldsd, etot);
}

__device__ __forceinline__ void gtile_counts(const GradArgs &G, uint32_t tile,
    const double *lw, double m, double invS, const double *toffh,
    const double *toffl, const UniformSeq &U, uint32_t *cnt, int gt,
    double *ldsd, uint32_t *ldsu, uint32_t &esum, uint32_t &zsum)
{
    const uint64_t i0 = tile_first(tile, gt);
    const uint64_t tend = (uint64_t)(tile + 1) * kTile < G.N
        ? (uint64_t)(tile + 1) * kTile
        : G.N;
    double W[kItems];
    dd loc{0.0, 0.0};
    load8d(lw, i0, G.N, W, 0.0);
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        W[k] = i < G.N ? exp(W[k] - m) : 0.0;
        loc = dd_add(loc, dd{W[k], 0.0});
    }
    dd total;
    const dd toff = group_excl_scan_dd(loc, gt, ldsd, total);
    const dd base = dd_add(dd{toffh[tile], toffl[tile]}, toff);
    // boundary counts shared with the neighbouring tiles; the counts are
    // made non-decreasing inside [clo, chi] (as in tile_counts_l), so the
    // replication numbers are never negative whatever the rounding
    const uint32_t clo = tile == 0 ? 0u : U.count_fast(toffh[tile] * invS);
    const uint32_t chi =
        tend == G.N ? U.M : U.count_fast(toffh[tile + 1] * invS);
    uint32_t c[kItems];
    dd run{0.0, 0.0};
    uint32_t cmax = clo;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        run = dd_add(run, dd{W[k], 0.0});
        c[k] = 0;
        if (i < G.N) {
            const uint32_t v = i == tend - 1
                ? chi
                : U.count_fast(dd_add(base, run).hi * invS);
            cmax = cmax > clamp_count(v, clo, chi) ? cmax
                                                   : clamp_count(v, clo, chi);
            c[k] = cmax;
        }
    }
    const uint32_t before = group_excl_max_u(cmax, clo, gt, ldsu);
    uint32_t prev = before;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < G.N) {
            c[k] = c[k] < before ? before : c[k];
            cnt[i] = c[k];
        }
    }
    uint32_t es = 0, zs = 0;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < G.N) {
            const uint32_t rep = rep_from_counts(c[k], prev);
            es += extra_copies(rep);
            zs += rep == 0 ? 1 : 0;
        }
        prev = c[k];
    }
    uint32_t te, tz;
    (void)group_excl_scan_u(es, gt, ldsu, te);
    (void)group_excl_scan_u(zs, gt, ldsu + 4, tz);
    esum = te;
    zsum = tz;
}

// tile_index on a slot: E, par (own index; zero slots are filled by
// tile_scatter_ranks) and the positions of the zero slots by rank, which the
// adjoint's gather over children also uses
__device__ __forceinline__ void gtile_index(const GradArgs &G, uint32_t tile,
    const uint32_t *cnt, uint32_t eoff, uint32_t zoff, const SlotPtr &sl,
    int gt, uint32_t *ldsu)
{
    const uint64_t i0 = tile_first(tile, gt);
    uint32_t rep[kItems];
    uint32_t es = 0, zs = 0;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        rep[k] = 0;
        if (i < G.N) {
            const uint32_t p = i == 0 ? 0u : cnt[i - 1];
            rep[k] = rep_from_counts(cnt[i], p);
            es += extra_copies(rep[k]);
            zs += rep[k] == 0 ? 1 : 0;
        }
    }
    uint32_t te, tz;
    uint32_t ex = group_excl_scan_u(es, gt, ldsu, te) + eoff;
    uint32_t zx = group_excl_scan_u(zs, gt, ldsu + 4, tz) + zoff;
    for (int k = 0; k < kItems; ++k) {
        const uint64_t i = i0 + k;
        if (i < G.N) {
            ex += extra_copies(rep[k]);
            sl.E[i] = ex;
            sl.par[i] = (uint32_t)i; // zero slots: overwritten by the scatter
            if (rep[k] == 0) {
                sl.zpos[zx] = (uint32_t)i;
                ++zx;
            }
        }
    }
}

} // namespace mckl_hip
