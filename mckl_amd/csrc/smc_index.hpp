// smc_index.hpp -- the index arithmetic of the resampling kernels that
// decides which memory they touch: cumulative counts confined to a tile's
// boundary counts, replication numbers, the range of extra-copy ranks a tile
// deals and the source of each rank.  Kept apart from the kernels (which add
// the workgroup scans and LDS staging around them) and __host__ __device__,
// so the same functions run on the CPU over adversarial inputs.
//
// Invariants the kernels rely on (algorithm/resample.hpp:374-466 for the
// replication vector and the stable rep -> index transform):
//   * a tile's counts are non-decreasing and inside [clo, chi], the counts at
//     its two boundaries, which its neighbours compute identically; so
//     rep_i = cnt_i - cnt_{i-1} >= 0 and the reps sum to M;
//   * a tile's ranks lie in [E_{t0-1}, min(E_{t0+n-1}, limit)), limit the
//     number of ranks that have a destination, so every rank reads an entry
//     of the zero-slot array and the loop is bounded even for a vector that
//     is not a replication vector;
//   * the source of a rank is a particle of the tile (< n).
#pragma once

#include <cmath>
#include <cstdint>

namespace mckl_hip {

// a particle's count confined to its tile's boundary counts
__host__ __device__ inline uint32_t clamp_count(
    uint32_t v, uint32_t clo, uint32_t chi)
{
    return v < clo ? clo : (v > chi ? chi : v);
}

// replication number of a particle from its cumulative count and the one
// before it (non-negative when the counts are non-decreasing)
__host__ __device__ inline uint32_t rep_from_counts(uint32_t c, uint32_t prev)
{
    return c - prev;
}

// extra copies of a source with replication number r
__host__ __device__ inline uint32_t extra_copies(uint32_t r)
{
    return r > 1 ? r - 1 : 0u;
}

// end of a tile's rank range: its inclusive extra-copy prefix, kept at or
// below `limit` (the ranks that have a destination)
__host__ __device__ inline uint32_t rank_end(uint32_t elast, uint64_t limit)
{
    return (uint64_t)elast < limit ? elast : (uint32_t)limit;
}

// source of rank r within a tile: the first j < n with Es[j] > r, for a
// non-decreasing Es whose last entry exceeds r (a rank below rank_end)
__host__ __device__ inline uint32_t rank_source(
    const uint32_t *Es, uint32_t n, uint32_t r)
{
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (Es[mid] > r)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// ---- counts of resampling uniforms below a cumulative weight known to
// within a margin ----
// The samplers evaluate a particle's cumulative weight S in double (Sf)
// with a bound mg on |S - Sf|, S the double-double prefix rounded once.
// When no uniform can lie within mg of Sf the count #{j < M : u_j < S} is
// the same for every S in [Sf - mg, Sf + mg] and follows from the uniforms'
// closed form; otherwise these return false and the caller counts against
// the exact S.  With unit roundoff u = 2^-53 and x = fl(fl(Sf - u0d) M):
//   systematic u_j = fl(fl(j delta) + u0d), delta = fl(1/M), u0d < delta:
//     |u_j - (j/M + u0d)| <= 3u (+ O(u^2)) and |x - (Sf - u0d) M| <= 2uM,
//     so u_j < S <=> j < x + th with |th| <= (mg + 5u) M: if x is farther
//     than D = (mg + 8u) M from every integer the count is
//     clamp(ceil(x), 0, M);
//   stratified u_j = fl(delta fl(U_j + j)), U_j in [0, 1), x = fl(Sf M):
//     u_j <= (j+1)/M + 3u and u_j >= j/M - 3u, so with x farther than D
//     from every integer every j < floor(x) counts, no j > floor(x) does,
//     and j = floor(x) counts when its own u_j (evaluated by the caller)
//     lies below Sf - mg, not when it lies at or above Sf + mg.
constexpr double kUnitRoundoff = 1.1102230246251565e-16; // 2^-53

__host__ __device__ inline bool count_sys_guarded(
    double Sf, double mg, double u0d, uint32_t M, uint32_t &c)
{
    const double x = (Sf - u0d) * (double)M;
    const double D = (mg + 8.0 * kUnitRoundoff) * (double)M;
    const double r = floor(x + 0.5);
    if (!(fabs(x - r) > D)) // also catches NaN
        return false;
    const double cx = ceil(x);
    c = cx <= 0.0 ? 0u : (cx >= (double)M ? M : (uint32_t)cx);
    return true;
}

// stratified: the index j = floor(x) whose own uniform decides (j < M), or
// the count itself when every uniform lies on one side (all counts: M).
// Returns 0 when the margin does not hold, 1 when c is the count, 2 when
// the caller must compare u_c with Sf -/+ mg (count c + 1 below, c above).
__host__ __device__ inline int count_strat_guarded(
    double Sf, double mg, uint32_t M, uint32_t &c)
{
    const double x = Sf * (double)M;
    const double D = (mg + 8.0 * kUnitRoundoff) * (double)M;
    const double r = floor(x + 0.5);
    if (!(fabs(x - r) > D))
        return 0;
    const double fx = floor(x);
    if (fx < 0.0) {
        c = 0;
        return 1;
    }
    if (fx >= (double)M) {
        c = M;
        return 1;
    }
    c = (uint32_t)fx;
    return 2;
}

} // namespace mckl_hip
