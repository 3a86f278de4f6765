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

} // namespace mckl_hip
