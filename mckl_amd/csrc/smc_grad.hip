// smc_grad.hip -- the batched forward/adjoint kernel of the SMC
// log-likelihood gradient and its C ABI (the function differentiated and the
// device pieces are described in smc_grad.hpp).
#include "smc_grad.hpp"

namespace mckl_hip {

struct GLds {
    double tmax[kMaxTilesG];
    double ts[kMaxTilesG], ts2[kMaxTilesG];
    double tdh[kMaxTilesG], tdl[kMaxTilesG];
    double toh[kMaxTilesG + 1], tol[kMaxTilesG + 1];
    double tmh[kMaxTilesG], tml[kMaxTilesG]; // multinomial: log spacings
    uint32_t te[kMaxTilesG], tz[kMaxTilesG];
    uint32_t teo[kMaxTilesG + 1], tzo[kMaxTilesG + 1];
    uint32_t ti[kMaxTilesG];               // residual: integral parts
    double trh[kMaxTilesG], trl[kMaxTilesG]; // residual: residual totals
    double red[kMaxTilesG];
    double csum[kMaxTilesG];
    double acc[3 * 4];
    int point;
};

struct GradSlotBufs {
    double *ck, *seg, *xb, *lwb, *xpb, *cb;
    uint32_t *cnt;
    double *u; // multinomial: the step's sorted uniforms (else null)
};

__host__ __device__ constexpr bool is_residual_g(int scheme)
{
    return scheme == MCKL_HIP_RESIDUAL ||
        scheme == MCKL_HIP_RESIDUAL_STRATIFIED ||
        scheme == MCKL_HIP_RESIDUAL_SYSTEMATIC;
}

// the schemes that draw sorted (multinomial) uniforms
__host__ __device__ constexpr bool needs_sorted(int scheme)
{
    return scheme == MCKL_HIP_MULTINOMIAL || scheme == MCKL_HIP_RESIDUAL;
}

// kernel instantiations by the resampling code they contain (each only its
// own, which keeps the others' register use down)
constexpr int kGradBase = 0;      // stratified, systematic
constexpr int kGradMult = 1;      // multinomial
constexpr int kGradResid = 2;     // residual stratified, residual systematic
constexpr int kGradResidMult = 3; // residual multinomial

__host__ __device__ constexpr int grad_kind(int scheme)
{
    return scheme == MCKL_HIP_MULTINOMIAL ? kGradMult
        : scheme == MCKL_HIP_RESIDUAL     ? kGradResidMult
        : is_residual_g(scheme)           ? kGradResid
                                          : kGradBase;
}

__global__ void k_grad_reset(uint32_t *queue)
{
    if (threadIdx.x == 0)
        *queue = 0;
}

constexpr int kGradGroups = 2; // 512 threads: 4 groups (1024) would spill at 128 VGPRs

// modes
constexpr int kGradRun = 0;    // forward + adjoint
constexpr int kGradRecord = 1; // forward + adjoint, writing the tape
constexpr int kGradReplay = 2; // forward with the tape's ancestors/decisions

template <int G>
__device__ void copy_slot(const SlotPtr &d, const SlotPtr &s, uint32_t Np)
{
    for (uint32_t i = threadIdx.x; i < Np; i += G * kGroup) {
        d.x[i] = s.x[i];
        d.lw[i] = s.lw[i];
        d.par[i] = s.par[i];
        d.E[i] = s.E[i];
        d.zpos[i] = s.zpos[i];
    }
    if (threadIdx.x < 4)
        d.scal[threadIdx.x] = s.scal[threadIdx.x];
    __syncthreads();
}

// Multinomial: the step's sorted uniforms into u (the samplers'
// log-spacing tiles, tile_mterms / tile_muniforms).  Only the multinomial
// instantiation of the kernel contains it: compiled into the others it
// raises their register use past what fits without spills.  Every thread
// of the workgroup calls it.
template <int G>
__device__ __forceinline__ void grad_muniforms(const GradArgs &A,
    uint32_t t, double *u, GLds &S, const GroupLds &L, int grp, int gt,
    uint32_t M)
{
    const int nt = (int)A.ntiles;
    const int rounds = (nt + G - 1) / G;
    SmcArgs ma{};
    ma.N = A.N;
    ma.k0 = A.k0;
    ma.k1 = A.k1;
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        dd mt;
        tile_mterms(ma, (uint32_t)(k < nt ? k : nt), t, gt, L.d, mt, M);
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
            ma, (uint32_t)kk, t, dd{S.toh[kk], S.tol[kk]}, u, gt, L.d, M);
    }
    __syncthreads();
}

// Residual resampling (ResampleAlgorithm<U01Seq, true>, resample.hpp:
// 333-364, 515-593) on the FP64 weights W_i = exp(lw_i - m): the counts of
// the FP32 samplers' residual step (tile_residual_sums_e, the R uniforms of
// the base scheme, tile_rcounts_e), leaving the tiles' extra-copy and
// zero-slot totals in S.te / S.tz.  The sorted uniforms of residual
// multinomial need the kMult instantiation.  Every thread of the workgroup
// calls it.
template <int G, bool kMult>
__device__ __forceinline__ void grad_rcounts(const GradArgs &A, uint32_t t,
    const double *lw, double m, double invS, double *u, uint32_t *cnt,
    GLds &S, const GroupLds &L, int grp, int gt)
{
    const int nt = (int)A.ntiles;
    const int rounds = (nt + G - 1) / G;
    SmcArgs ra{};
    ra.N = A.N;
    ra.k0 = A.k0;
    ra.k1 = A.k1;
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        const uint32_t kk = (uint32_t)(k < nt ? k : nt);
        double e[kItems];
        load8d(lw, tile_first(kk, gt), A.N, e, 0.0);
#pragma unroll
        for (int q = 0; q < kItems; ++q)
            e[q] = tile_first(kk, gt) + q < A.N ? exp(e[q] - m) : 0.0;
        uint32_t isum;
        dd rt;
        tile_residual_sums_e(ra, kk, e, invS, gt, L.d, L.u, isum, rt);
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
    if constexpr (kMult) {
        if (A.scheme == MCKL_HIP_RESIDUAL)
            grad_muniforms<G>(A, t, u, S, L, grp, gt, R);
    }
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
        const uint32_t kk = (uint32_t)(k < nt ? k : nt);
        uint32_t ioff = 0;
        for (uint32_t i = 0; i < kk; ++i)
            ioff += S.ti[i];
        double e[kItems];
        load8d(lw, tile_first(kk, gt), A.N, e, 0.0);
#pragma unroll
        for (int q = 0; q < kItems; ++q)
            e[q] = tile_first(kk, gt) + q < A.N ? exp(e[q] - m) : 0.0;
        uint32_t es, zs;
        tile_rcounts_e(ra, kk, e, invS, ioff, dd{S.toh[kk], S.tol[kk]},
            (int)kk < nt ? S.toh[kk + 1] : 0.0, invR, U, cnt, gt, L.d, L.u,
            es, zs);
        if (gt == 0 && k < nt) {
            S.te[k] = es;
            S.tz[k] = zs;
        }
        __syncthreads();
    }
}

// Multinomial: the step's sorted uniforms, the weights' tile offsets, then
// the counts of each tile against the uniforms (binary search), leaving
// the tiles' extra-copy and zero-slot totals in S.te / S.tz (multinomial
// instantiation only, as grad_muniforms).  Every thread of the workgroup
// calls it.
template <int G>
__device__ __forceinline__ void grad_mcounts(const GradArgs &A,
    uint32_t t, const double *lw, double m, double invS, double *u,
    uint32_t *cnt, GLds &S, const GroupLds &L, int grp, int gt)
{
    const int nt = (int)A.ntiles;
    const int rounds = (nt + G - 1) / G;
    grad_muniforms<G>(A, t, u, S, L, grp, gt, A.N); // uses S.toh/S.tol
    tiles_scan_dd(grp == 0, S.tdh, S.tdl, nt, S.toh, S.tol, gt, L.d);
    UniformSeq U;
    U.u = u;
    U.scheme = MCKL_HIP_MULTINOMIAL;
    U.k0 = A.k0;
    U.k1 = A.k1;
    U.t = t;
    U.M = A.N;
    U.delta = 1.0 / (double)A.N;
    U.u0d = 0.0;
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        uint32_t es, zs;
        gtile_counts(A, (uint32_t)(k < nt ? k : nt), lw, m, invS, S.toh,
            S.tol, U, cnt, gt, L.d, L.u, es, zs);
        if (gt == 0 && k < nt) {
            S.te[k] = es;
            S.tz[k] = zs;
        }
        __syncthreads();
    }
}

// One forward step t from slot `prev` to slot `cur`.  Returns the step's
// contribution to F.  `rt` is the tape row pointer set for record/replay.
// Kind: the instantiation (grad_kind; u: the step's sorted uniforms for the
// kinds that draw them).
template <int G, int Kind>
__device__ double fwd_step(const GradArgs &A, uint32_t t, const SlotPtr &prev,
    const SlotPtr &cur, uint32_t *cnt, double *u, int mode,
    const double *lwn_row,
    const uint32_t *par_row, int32_t *tape_r, double *tape_lwn_out,
    uint32_t *tape_par_out, GLds &S, const GroupLds &L, int grp, int gt,
    uint32_t *Es)
{
    const int nt = (int)A.ntiles;
    const int rounds = (nt + G - 1) / G;
    const int Rprev = prev.scal[3] != 0.0;
    const double mprev = prev.scal[0];
    const bool replay = mode == kGradReplay;
    // move + weight; with resampled parents also LSE(carry) by tile
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        const uint32_t kk = (uint32_t)(k < nt ? k : nt);
        const double mt = gtile_move(A, t, kk, prev, cur, Rprev, mprev,
            replay ? lwn_row : nullptr, replay ? par_row : nullptr, gt, L.d);
        double cs = 0.0, dummy = 0.0;
        if (Rprev) {
            const uint64_t i0 = tile_first(kk, gt);
            for (int q = 0; q < kItems; ++q) {
                const uint64_t i = i0 + q;
                if (i < A.N) {
                    double carry = 0.0;
                    if (replay) {
                        const uint32_t p = par_row[i];
                        carry = (prev.lw[p] - prev.scal[2]) - lwn_row[p];
                    }
                    cs += exp(carry);
                }
            }
            group_sum2(cs, dummy, gt, L.d);
        }
        if (gt == 0 && k < nt) {
            S.tmax[k] = mt;
            S.csum[k] = cs;
        }
        __syncthreads();
    }
    double m = -__builtin_huge_val();
    for (int i = 0; i < nt; ++i)
        m = m < S.tmax[i] ? S.tmax[i] : m;
    for (int r = 0; r < rounds; ++r) {
        const int k = r * G + grp;
        double s, s2;
        dd et;
        gtile_sums(A, (uint32_t)(k < nt ? k : nt), cur.lw, m, gt, L.d, s, s2, et);
        if (gt == 0 && k < nt) {
            S.ts[k] = s;
            S.ts2[k] = s2;
            S.tdh[k] = et.hi;
            S.tdl[k] = et.lo;
        }
        __syncthreads();
    }
    double rs, rs2;
    tiles_sum2(S.ts, S.ts2, nt, gt, L.d, rs, rs2);
    double C;
    if (Rprev) {
        C = 0.0;
        for (int i = 0; i < nt; ++i)
            C += S.csum[i];
    } else {
        C = prev.scal[1];
    }
    const double lse = m + log(rs);
    const double inc = lse - log(C);
    const double ess = rs * rs / rs2;
    int R;
    if (replay)
        R = tape_r[t];
    else
        R = ess < A.tauN ? 1 : 0;
    if (threadIdx.x == 0) {
        cur.scal[0] = m;
        cur.scal[1] = rs;
        cur.scal[2] = lse;
        cur.scal[3] = (double)R;
        if (mode == kGradRecord)
            tape_r[t] = R;
    }
    __syncthreads();
    if (R && !replay) {
        const double invS = 1.0 / rs;
        if constexpr (Kind == kGradResid || Kind == kGradResidMult) {
            grad_rcounts<G, Kind == kGradResidMult>(A, t, cur.lw, m, invS, u,
                cnt, S, L, grp, gt);
        } else if constexpr (Kind == kGradMult) {
            grad_mcounts<G>(A, t, cur.lw, m, invS, u, cnt, S, L, grp, gt);
        } else {
            tiles_scan_dd(grp == 0, S.tdh, S.tdl, nt, S.toh, S.tol, gt, L.d);
            UniformSeq U;
            U.u = nullptr;
            U.scheme = A.scheme;
            U.k0 = A.k0;
            U.k1 = A.k1;
            U.t = t;
            U.M = A.N;
            U.delta = 1.0 / (double)A.N;
            U.u0d = A.scheme == MCKL_HIP_SYSTEMATIC
                ? smc_uniform(A.k0, A.k1, t, 0) * U.delta
                : 0.0;
            for (int r = 0; r < rounds; ++r) {
                const int k = r * G + grp;
                uint32_t es, zs;
                gtile_counts(A, (uint32_t)(k < nt ? k : nt), cur.lw, m, invS,
                    S.toh, S.tol, U, cnt, gt, L.d, L.u, es, zs);
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
            gtile_index(A, (uint32_t)kk, cnt, S.teo[kk], S.tzo[kk], cur, gt, L.u);
            __syncthreads();
        }
        // zero slots take their sources (resolved parents, no search later)
        SmcArgs sa{};
        sa.N = A.N;
        for (int r = 0; r < rounds; ++r) {
            const int k = r * G + grp;
            tile_scatter_ranks(sa, (uint32_t)(k < nt ? k : nt), cur.E,
                cur.zpos, cur.par, gt, Es);
        }
        if (mode == kGradRecord) {
            for (uint32_t i = threadIdx.x; i < A.N; i += G * kGroup) {
                tape_par_out[i] = cur.par[i];
                tape_lwn_out[i] = cur.lw[i] - lse;
            }
            __syncthreads();
        }
    }
    return inc;
}

template <int G, int Kind>
__global__ __launch_bounds__(G *kGroup) void k_grad(const GradArgs *args,
    uint32_t npoints, uint32_t *queue, GradSlotBufs B, size_t slot_bytes,
    size_t ck_stride, size_t seg_stride, size_t vec_stride, int mode,
    double *tape_lwn, uint32_t *tape_par, int32_t *tape_r, double *logz_out,
    double *grad_out)
{
    __shared__ __align__(16) unsigned char scratch[G * kScratch];
    __shared__ GLds S;
    __shared__ uint32_t Es[G * kTile]; // tile_scatter_ranks, one per group
    const int grp = threadIdx.x / kGroup;
    const int gt = threadIdx.x % kGroup;
    const GroupLds L = group_lds(scratch + grp * kScratch);
    const size_t slot = blockIdx.x;
    double *ck = B.ck + slot * ck_stride;
    double *seg = B.seg + slot * seg_stride;
    double *xb = B.xb + slot * vec_stride;
    double *lwb = B.lwb + slot * vec_stride;
    double *xpb = B.xpb + slot * vec_stride;
    double *cb = B.cb + slot * vec_stride;
    uint32_t *cnt = B.cnt + slot * vec_stride;
    double *ub = B.u ? B.u + slot * vec_stride : nullptr;
    for (;;) {
        if (threadIdx.x == 0) {
            const uint32_t q = atomicAdd(queue, 1u);
            S.point = q < npoints ? (int)q : -1;
        }
        __syncthreads();
        const int p = S.point;
        __syncthreads();
        if (p < 0)
            return;
        const GradArgs A = args[p];
        const int nt = (int)A.ntiles;
        const int rounds = (nt + G - 1) / G;
        const uint32_t Np = A.ntiles * kTile;
        const size_t rowN = A.N;
        double *tl = tape_lwn ? tape_lwn + (size_t)p * (A.T + 1) * rowN : nullptr;
        uint32_t *tp = tape_par ? tape_par + (size_t)p * (A.T + 1) * rowN : nullptr;
        int32_t *tr = tape_r ? tape_r + (size_t)p * (A.T + 1) : nullptr;

        // ---------------- forward ----------------
        SlotPtr s0 = slot_at(seg, slot_bytes, 0, Np);
        SlotPtr s1 = slot_at(seg, slot_bytes, 1, Np);
        for (uint32_t i = threadIdx.x; i < A.N; i += G * kGroup) {
            s0.x[i] = A.sd0 * eps_at(A, i, 0);
            s0.lw[i] = 0.0;
        }
        if (threadIdx.x == 0) {
            s0.scal[0] = 0.0;
            s0.scal[1] = (double)A.N;
            s0.scal[2] = log((double)A.N);
            s0.scal[3] = 0.0;
            if (mode == kGradRecord)
                tr[0] = 0;
        }
        __syncthreads();
        if (mode != kGradReplay)
            copy_slot<G>(slot_at(ck, slot_bytes, 0, Np), s0, Np);
        double F = 0.0;
        for (uint32_t t = 1; t <= A.T; ++t) {
            const SlotPtr &prev = (t & 1) ? s0 : s1;
            const SlotPtr &cur = (t & 1) ? s1 : s0;
            F += fwd_step<G, Kind>(A, t, prev, cur, cnt, ub, mode,
                tl ? tl + (size_t)(t - 1) * rowN : nullptr,
                tp ? tp + (size_t)(t - 1) * rowN : nullptr, tr,
                tl ? tl + (size_t)t * rowN : nullptr,
                tp ? tp + (size_t)t * rowN : nullptr, S, L, grp, gt,
                Es + grp * kTile);
            if (mode != kGradReplay && t % A.K == 0 && t < A.T)
                copy_slot<G>(slot_at(ck, slot_bytes, t / A.K, Np), cur, Np);
        }
        if (threadIdx.x == 0)
            logz_out[p] = F;
        __syncthreads();
        if (mode == kGradReplay)
            continue;

        // ---------------- backward ----------------
        for (uint32_t i = threadIdx.x; i < Np; i += G * kGroup) {
            xb[i] = 0.0;
            lwb[i] = 0.0;
        }
        __syncthreads();
        double ta = 0.0, tsg = 0.0, tc = 0.0;
        const uint32_t nseg = (A.T + A.K - 1) / A.K;
        for (int c = (int)nseg - 1; c >= 0; --c) {
            const uint32_t t0 = (uint32_t)c * A.K;
            const uint32_t t1 = t0 + A.K < A.T ? t0 + A.K : A.T;
            copy_slot<G>(slot_at(seg, slot_bytes, 0, Np),
                slot_at(ck, slot_bytes, (uint32_t)c, Np), Np);
            for (uint32_t t = t0 + 1; t <= t1; ++t)
                (void)fwd_step<G, Kind>(A, t,
                    slot_at(seg, slot_bytes, t - t0 - 1, Np),
                    slot_at(seg, slot_bytes, t - t0, Np), cnt, ub, kGradRun,
                    nullptr,
                    nullptr, nullptr, nullptr, nullptr, S, L, grp, gt,
                    Es + grp * kTile);
            for (uint32_t t = t1; t > t0; --t) {
                const SlotPtr cur = slot_at(seg, slot_bytes, t - t0, Np);
                const SlotPtr prev = slot_at(seg, slot_bytes, t - t0 - 1, Np);
                const int Rprev = prev.scal[3] != 0.0;
                const double lse_t = cur.scal[2], lse_p = prev.scal[2];
                const double y = A.y[t - 1];
                const double invN = 1.0 / (double)A.N;
                // phase 1: per child
                for (int r = 0; r < rounds; ++r) {
                    const int k = r * G + grp;
                    const uint32_t kk = (uint32_t)(k < nt ? k : nt);
                    const uint64_t i0 = tile_first(kk, gt);
                    double csum = 0.0, dummy = 0.0;
                    float ep[2] = {0.0f, 0.0f}; // the normal pair of i, i^1
                    if (i0 + kItems <= A.N) {
                        // a whole row: loads issued together, then the
                        // element loop's arithmetic in the same order
                        double lwt[kItems], lwp[kItems], lbv[kItems];
                        double xc[kItems], xbv[kItems], xpv[kItems];
                        float e[kItems];
                        uint32_t par[kItems];
                        load8d(cur.lw, i0, A.N, lwt, 0.0);
                        load8d(lwb, i0, A.N, lbv, 0.0);
                        load8d(cur.x, i0, A.N, xc, 0.0);
                        load8d(xb, i0, A.N, xbv, 0.0);
                        if (Rprev) {
                            load8(prev.par, i0, A.N, par, 0u);
#pragma unroll
                            for (int q = 0; q < kItems; ++q)
                                xpv[q] = prev.x[par[q]];
                        } else {
                            load8d(prev.lw, i0, A.N, lwp, 0.0);
                            load8d(prev.x, i0, A.N, xpv, 0.0);
                        }
#pragma unroll
                        for (int q = 0; q < kItems; q += 2)
                            smc_normal_pair(A.k0, A.k1, (i0 + q) >> 1, t,
                                e[q], e[q + 1]);
                        double o1[kItems], o2[kItems];
#pragma unroll
                        for (int q = 0; q < kItems; ++q) {
                            const double Wt = exp(lwt[q] - lse_t);
                            const double Wc =
                                Rprev ? invN : exp(lwp[q] - lse_p);
                            const double lwbar = lbv[q] + Wt;
                            const double cbar = lwbar - Wc;
                            double g, gx, gc;
                            gmodel(A, y, xc[q], g, gx, gc);
                            const double xbar = xbv[q] + lwbar * gx;
                            tc += lwbar * gc;
                            ta += xbar * xpv[q];
                            tsg += xbar * (double)e[q];
                            o1[q] = A.a * xbar;
                            o2[q] = cbar;
                            if (Rprev)
                                csum += cbar;
                        }
                        store8d(Rprev ? xpb : xb, i0, A.N, o1);
                        store8d(Rprev ? cb : lwb, i0, A.N, o2);
                    } else
                    for (int q = 0; q < kItems; ++q) {
                        const uint64_t i = i0 + q;
                        if (i >= A.N)
                            break;
                        if ((q & 1) == 0)
                            smc_normal_pair(A.k0, A.k1, i >> 1, t, ep[0], ep[1]);
                        const double Wt = exp(cur.lw[i] - lse_t);
                        const double Wc = Rprev ? invN : exp(prev.lw[i] - lse_p);
                        const double lwbar = lwb[i] + Wt;
                        const double cbar = lwbar - Wc;
                        double g, gx, gc;
                        gmodel(A, y, cur.x[i], g, gx, gc);
                        const double xbar = xb[i] + lwbar * gx;
                        tc += lwbar * gc;
                        const uint32_t par =
                            Rprev ? prev.par[i]
                                  : (uint32_t)i;
                        ta += xbar * prev.x[par];
                        tsg += xbar * (double)ep[q & 1];
                        const double xpbar = A.a * xbar;
                        if (Rprev) {
                            xpb[i] = xpbar;
                            cb[i] = cbar;
                            csum += cbar;
                        } else {
                            xb[i] = xpbar;
                            lwb[i] = cbar;
                        }
                    }
                    if (Rprev) {
                        group_sum2(csum, dummy, gt, L.d);
                        if (gt == 0 && k < nt)
                            S.red[k] = csum;
                    }
                    __syncthreads();
                }
                if (!Rprev)
                    continue;
                // phase 2: per parent, sum over its children (own slot and
                // the zero slots holding its extra copies, in rank order)
                double total = 0.0;
                for (int i = 0; i < nt; ++i)
                    total += S.red[i];
                for (int r = 0; r < rounds; ++r) {
                    const int k = r * G + grp;
                    if (k >= nt)
                        continue;
                    const uint64_t i0 = tile_first((uint32_t)k, gt);
                    if (i0 >= A.N)
                        continue;
                    // the row's contiguous values loaded together
                    uint32_t parv[kItems], Ev[kItems];
                    double xpv[kItems], cbv[kItems], lpv[kItems];
                    load8(prev.par, i0, A.N, parv, 0u);
                    load8(prev.E, i0, A.N, Ev, 0u);
                    load8d(xpb, i0, A.N, xpv, 0.0);
                    load8d(cb, i0, A.N, cbv, 0.0);
                    load8d(prev.lw, i0, A.N, lpv, 0.0);
                    uint32_t Eprev = i0 == 0 ? 0u : prev.E[i0 - 1];
                    double xo[kItems], lo[kItems];
#pragma unroll
                    for (int q = 0; q < kItems; ++q) {
                        const uint64_t j = i0 + q;
                        const bool own = parv[q] == (uint32_t)j;
                        double xs = own ? xpv[q] : 0.0;
                        double cs = own ? cbv[q] : 0.0;
                        // ranks of j's extra copies, kept inside zpos
                        const uint32_t Ej = rank_end(Ev[q], A.N);
                        const uint32_t Ejm = Eprev;
                        Eprev = Ev[q];
                        for (uint32_t rk = Ejm; rk < Ej; ++rk) {
                            const uint32_t ch = prev.zpos[rk];
                            if (ch < A.N) {
                                xs += xpb[ch];
                                cs += cb[ch];
                            }
                        }
                        const double Wp = exp(lpv[q] - lse_p);
                        xo[q] = xs;
                        lo[q] = cs - Wp * total;
                    }
                    store8d(xb, i0, A.N, xo);
                    store8d(lwb, i0, A.N, lo);
                }
                __syncthreads();
            }
        }
        // t = 0: x_0 = sd0 eps_0
        for (int r = 0; r < rounds; ++r) {
            const int k = r * G + grp;
            if (k >= nt)
                continue;
            const uint64_t i0 = tile_first((uint32_t)k, gt);
            for (int q = 0; q < kItems; ++q) {
                const uint64_t i = i0 + q;
                if (i >= A.N)
                    break;
                const double e0 = eps_at(A, i, 0);
                ta += xb[i] * e0 * A.dsd0_a;
                tsg += xb[i] * e0 * A.dsd0_s;
            }
        }
        double d0 = 0.0;
        group_sum2(ta, tsg, gt, L.d);
        group_sum2(tc, d0, gt, L.d);
        if (gt == 0) {
            S.acc[3 * grp + 0] = ta;
            S.acc[3 * grp + 1] = tsg;
            S.acc[3 * grp + 2] = tc;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double g3[3] = {0.0, 0.0, 0.0};
            for (int gg = 0; gg < G; ++gg)
                for (int j = 0; j < 3; ++j)
                    g3[j] += S.acc[3 * gg + j];
            for (int j = 0; j < 3; ++j)
                grad_out[3 * (size_t)p + j] = g3[j];
        }
        __syncthreads();
    }
}

} // namespace mckl_hip

using namespace mckl_hip;

struct mckl_hip_smc_grad {
    int model = 0;
    size_t nbatch = 0, N = 0, T = 0, K = 0;
    int scheme = 0;
    double threshold = 0.5;
    std::vector<uint64_t> seeds;
    int y_per_point = 0;
    double *y = nullptr; // device observations
    GradArgs *args = nullptr;
    uint32_t *queue = nullptr;
    GradSlotBufs B{};
    size_t slot_bytes = 0, ck_stride = 0, seg_stride = 0, vec_stride = 0;
    int nslots = 0;
    double *logz = nullptr, *grad = nullptr;
    double *tape_lwn = nullptr;
    uint32_t *tape_par = nullptr;
    int32_t *tape_r = nullptr;
    int have_tape = 0;
};

static void grad_free(mckl_hip_smc_grad *g)
{
    void *ptrs[] = {g->y, g->args, g->queue, g->B.ck, g->B.seg, g->B.xb,
        g->B.lwb, g->B.xpb, g->B.cb, g->B.cnt, g->B.u, g->logz, g->grad,
        g->tape_lwn,
        g->tape_par, g->tape_r};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    delete g;
}

static int grad_args(mckl_hip_smc_grad *g, const double *h_theta,
    std::vector<GradArgs> &out)
{
    out.resize(g->nbatch);
    const double l2pi = std::log(2.0 * 3.141592653589793);
    for (size_t p = 0; p < g->nbatch; ++p) {
        const double *th = h_theta + 3 * p;
        if (!(th[1] > 0) || !(th[2] > 0))
            return set_error(MCKL_HIP_EINVAL, "scale parameters must be > 0");
        if (g->model == 2 && !(std::fabs(th[0]) < 1))
            return set_error(MCKL_HIP_EINVAL, "SV requires |phi| < 1");
        GradArgs &A = out[p];
        A.model = g->model;
        A.a = th[0];
        A.s = th[1];
        A.c = th[2];
        if (g->model == 1) {
            A.sd0 = 1.0;
            A.dsd0_a = 0.0;
            A.dsd0_s = 0.0;
        } else {
            const double v = 1.0 - th[0] * th[0];
            A.sd0 = th[1] / std::sqrt(v);
            A.dsd0_a = th[1] * th[0] / (v * std::sqrt(v));
            A.dsd0_s = 1.0 / std::sqrt(v);
        }
        A.cst = -std::log(th[2]) - 0.5 * l2pi;
        A.k0 = (uint32_t)g->seeds[p];
        A.k1 = (uint32_t)(g->seeds[p] >> 32);
        A.N = (uint32_t)g->N;
        A.ntiles = (uint32_t)ceil_div(g->N, kTile);
        A.T = (uint32_t)g->T;
        A.K = (uint32_t)g->K;
        A.tauN = g->threshold * (double)g->N;
        A.scheme = g->scheme;
        A.y = g->y + (g->y_per_point ? p * g->T : 0);
    }
    return MCKL_HIP_OK;
}

static int grad_launch(mckl_hip_smc_grad *g, const double *h_theta, int mode,
    hipStream_t st)
{
    std::vector<GradArgs> args;
    int rc = grad_args(g, h_theta, args);
    if (rc)
        return rc;
    MCKL_HIP_TRY(hipMemcpyAsync(g->args, args.data(),
        args.size() * sizeof(GradArgs), hipMemcpyHostToDevice, st));
    k_grad_reset<<<1, 64, 0, st>>>(g->queue);
    const bool tape = mode != kGradRun;
    const int kind = grad_kind(g->scheme);
    auto kern = kind == kGradMult ? k_grad<kGradGroups, kGradMult>
        : kind == kGradResid      ? k_grad<kGradGroups, kGradResid>
        : kind == kGradResidMult  ? k_grad<kGradGroups, kGradResidMult>
                                  : k_grad<kGradGroups, kGradBase>;
    kern<<<g->nslots, kGradGroups * kGroup, 0, st>>>(g->args,
        (uint32_t)g->nbatch, g->queue, g->B, g->slot_bytes, g->ck_stride,
        g->seg_stride, g->vec_stride, mode, tape ? g->tape_lwn : nullptr,
        tape ? g->tape_par : nullptr, tape ? g->tape_r : nullptr, g->logz,
        g->grad);
    MCKL_HIP_CHECK_LAUNCH("k_grad");
    MCKL_HIP_TRY(hipStreamSynchronize(st));
    return MCKL_HIP_OK;
}

extern "C" {

int mckl_hip_smc_grad_create(int model, size_t nbatch, size_t N, size_t T,
    const double *h_y, int y_per_point, const uint64_t *h_seed,
    double threshold, int scheme, size_t K, int tape,
    mckl_hip_smc_grad **out)
{
    if (!out || !h_y || !h_seed)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    *out = nullptr;
    if (model != 1 && model != 2)
        return set_error(MCKL_HIP_EINVAL, "unknown model %d", model);
    if (nbatch == 0 || N == 0 || T == 0 || K == 0)
        return set_error(MCKL_HIP_EINVAL, "sizes must be > 0");
    if (N > (size_t)kMaxTilesG * kTile)
        return set_error(MCKL_HIP_EINVAL,
            "the gradient sampler holds at most %d particles per point",
            kMaxTilesG * kTile);
    if (scheme < MCKL_HIP_MULTINOMIAL || scheme > MCKL_HIP_RESIDUAL_SYSTEMATIC)
        return set_error(MCKL_HIP_EINVAL, "unknown resampling scheme %d",
            scheme);
    if (std::isnan(threshold))
        return set_error(MCKL_HIP_EINVAL, "threshold is NaN");
    mckl_hip_smc_grad *g = new mckl_hip_smc_grad();
    g->model = model;
    g->nbatch = nbatch;
    g->N = N;
    g->T = T;
    g->K = K;
    g->scheme = scheme;
    g->threshold = threshold;
    g->seeds.assign(h_seed, h_seed + nbatch);
    g->y_per_point = y_per_point;
    const size_t Np = ceil_div(N, kTile) * kTile;
    g->slot_bytes = (Np * 28 + 32 + 255) / 256 * 256;
    const size_t nck = T / K + 1;
    g->ck_stride = nck * g->slot_bytes / sizeof(double);
    g->seg_stride = (K + 1) * g->slot_bytes / sizeof(double);
    g->vec_stride = Np;
    int dev = 0, ncu = 0;
    MCKL_HIP_TRY(hipGetDevice(&dev));
    MCKL_HIP_TRY(hipDeviceGetAttribute(
        &ncu, hipDeviceAttributeMultiprocessorCount, dev));
    size_t freeb = 0, totalb = 0;
    MCKL_HIP_TRY(hipMemGetInfo(&freeb, &totalb));
    // per resident slot: checkpoints, segment, four adjoint vectors, the
    // counts and, for multinomial and residual multinomial, the sorted
    // uniforms
    const size_t per_slot = (nck + K + 1) * g->slot_bytes + Np * 8 * 4 +
        Np * 4 + (needs_sorted(scheme) ? Np * 8 : 0);
    size_t nslots = std::min<size_t>(nbatch, (size_t)ncu);
    while (nslots > 1 && nslots * per_slot > freeb / 2)
        nslots /= 2;
    if (nslots * per_slot > freeb) {
        grad_free(g);
        return set_error(MCKL_HIP_ENOMEM,
            "smc_grad: one slot needs %zu bytes, %zu free", per_slot, freeb);
    }
    if (tape) {
        // (T + 1) N entries per point of 8 (log weights) + 4 (parents) bytes
        const uint64_t per_point = ((uint64_t)T + 1) * (uint64_t)N;
        if (per_point > UINT64_MAX / 12 / nbatch ||
            per_point * 12 * nbatch > (uint64_t)freeb / 2) {
            grad_free(g);
            return set_error(MCKL_HIP_EINVAL,
                "smc_grad: the ancestor tape (%zu points x %llu entries x 12 "
                "bytes) does not fit in half of the %zu free bytes", nbatch,
                (unsigned long long)per_point, freeb);
        }
    }
    g->nslots = (int)nslots;
    int rc = MCKL_HIP_OK;
#define MCKL_ALLOC(p, n)                                                      \
    do {                                                                      \
        rc = dev_alloc(&(p), (n));                                          \
        if (rc) {                                                             \
            grad_free(g);                                                     \
            return rc;                                                        \
        }                                                                     \
    } while (0)
    const size_t ny = y_per_point ? nbatch * T : T;
    MCKL_ALLOC(g->y, ny);
    MCKL_ALLOC(g->args, nbatch);
    MCKL_ALLOC(g->queue, 1);
    MCKL_ALLOC(g->B.ck, nslots * g->ck_stride);
    MCKL_ALLOC(g->B.seg, nslots * g->seg_stride);
    MCKL_ALLOC(g->B.xb, nslots * Np);
    MCKL_ALLOC(g->B.lwb, nslots * Np);
    MCKL_ALLOC(g->B.xpb, nslots * Np);
    MCKL_ALLOC(g->B.cb, nslots * Np);
    MCKL_ALLOC(g->B.cnt, nslots * Np);
    if (needs_sorted(scheme))
        MCKL_ALLOC(g->B.u, nslots * Np);
    MCKL_ALLOC(g->logz, nbatch);
    MCKL_ALLOC(g->grad, 3 * nbatch);
    if (tape) {
        MCKL_ALLOC(g->tape_lwn, nbatch * (T + 1) * N);
        MCKL_ALLOC(g->tape_par, nbatch * (T + 1) * N);
        MCKL_ALLOC(g->tape_r, nbatch * (T + 1));
    }
#undef MCKL_ALLOC
    hipError_t err = hipMemcpy(g->y, h_y, ny * sizeof(double),
        hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        rc = hip_status(err, "mckl_hip_smc_grad_create");
        grad_free(g);
        return rc;
    }
    *out = g;
    return MCKL_HIP_OK;
}

int mckl_hip_smc_grad_run(mckl_hip_smc_grad *g, const double *h_theta,
    double *h_loglik, double *h_grad, void *stream)
{
    if (!g || !h_theta)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    hipStream_t st = as_stream(stream);
    const int mode = g->tape_lwn ? kGradRecord : kGradRun;
    int rc = grad_launch(g, h_theta, mode, st);
    if (rc)
        return rc;
    g->have_tape = g->tape_lwn != nullptr;
    // grad_launch synchronised st: read on it (not on the null stream)
    if (h_loglik)
        MCKL_HIP_TRY(hipMemcpyWithStream(h_loglik, g->logz,
            g->nbatch * sizeof(double), hipMemcpyDeviceToHost, st));
    if (h_grad)
        MCKL_HIP_TRY(hipMemcpyWithStream(h_grad, g->grad,
            3 * g->nbatch * sizeof(double), hipMemcpyDeviceToHost, st));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_grad_replay(mckl_hip_smc_grad *g, const double *h_theta,
    double *h_value, void *stream)
{
    if (!g || !h_theta || !h_value)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (!g->have_tape)
        return set_error(MCKL_HIP_EINVAL,
            "replay needs a tape: create with tape = 1 and run first");
    int rc = grad_launch(g, h_theta, kGradReplay, as_stream(stream));
    if (rc)
        return rc;
    MCKL_HIP_TRY(hipMemcpyWithStream(h_value, g->logz,
        g->nbatch * sizeof(double), hipMemcpyDeviceToHost, as_stream(stream)));
    return MCKL_HIP_OK;
}

int mckl_hip_smc_grad_destroy(mckl_hip_smc_grad *g)
{
    if (g)
        grad_free(g);
    return MCKL_HIP_OK;
}

int mckl_hip_smc_loglik_grad(int model, size_t nbatch, const double *h_theta,
    size_t N, size_t T, const double *h_y, int y_per_point,
    const uint64_t *h_seed, double threshold, int scheme, size_t K,
    double *h_loglik, double *h_grad, void *stream)
{
    mckl_hip_smc_grad *g = nullptr;
    int rc = mckl_hip_smc_grad_create(model, nbatch, N, T, h_y, y_per_point,
        h_seed, threshold, scheme, K, 0, &g);
    if (rc)
        return rc;
    rc = mckl_hip_smc_grad_run(g, h_theta, h_loglik, h_grad, stream);
    grad_free(g);
    return rc;
}

} // extern "C"
