// The resampling kernels' index arithmetic (mckl_amd/csrc/smc_index.hpp)
// run on the CPU over a whole particle array, tile by tile, with the
// workgroup scans done sequentially: the stages of the fused samplers'
// tile_counts_l (clamp and monotone counts), tile_index_c (replication
// numbers, extra-copy prefix E, zero-slot positions zpos) and
// tile_scatter_ranks (each extra-copy rank's source, written to its zero
// slot).  Every array access the kernels make is checked against the
// array's length.  `guarded` = 0 replays the form before the count clamp and
// the rank bound (raw counts, ranks up to E of the tile's last particle),
// to show which accesses the guards keep inside the arrays.  Test
// infrastructure only.
#define __host__
#define __device__
#include "../../mckl_amd/csrc/smc_index.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

using namespace mckl_hip;

namespace {
constexpr uint32_t kGroup = 256, kItems = 8, kTile = kGroup * kItems;
constexpr uint32_t kStale = 0xFFFFFFFFu; // an entry no stage wrote
}

extern "C" {

// N particles (M = N), raw[i]: the count of uniforms below particle i's
// cumulative weight as computed from its own prefix; clo[k], chi[k]: tile
// k's boundary counts (the device computes chi[k] and clo[k+1] from one
// value).  Outputs idx[N], rep[N]; returns the number of out-of-range
// accesses attempted.
uint64_t rs_guard_run(uint64_t N, const uint32_t *raw, const uint32_t *clo,
    const uint32_t *chi, int guarded, uint32_t *idx, uint32_t *rep)
{
    uint64_t bad = 0;
    const uint64_t nt = (N + kTile - 1) / kTile;
    std::vector<uint32_t> cnt(N), E(N), zpos(N, kStale);
    // tile_counts_l: per thread a running max of the clamped counts, then
    // the exclusive max-scan over the group's threads
    for (uint64_t k = 0; k < nt; ++k) {
        const uint64_t t0 = k * kTile;
        const uint64_t tend = t0 + kTile < N ? t0 + kTile : N;
        const uint32_t lo = k == 0 ? 0u : clo[k];
        const uint32_t hi = k + 1 == nt ? (uint32_t)N : chi[k];
        std::vector<uint32_t> cmax(kGroup);
        for (uint32_t gt = 0; gt < kGroup; ++gt) {
            uint32_t m = lo;
            for (uint32_t q = 0; q < kItems; ++q) {
                const uint64_t i = t0 + gt * kItems + q;
                if (i >= N)
                    continue;
                uint32_t v = i == tend - 1 ? hi : raw[i];
                if (guarded) {
                    v = clamp_count(v, lo, hi);
                    m = m > v ? m : v;
                    cnt[i] = m;
                } else {
                    cnt[i] = v;
                }
            }
            cmax[gt] = m;
        }
        if (guarded) {
            uint32_t before = lo;
            for (uint32_t gt = 0; gt < kGroup; ++gt) {
                for (uint32_t q = 0; q < kItems; ++q) {
                    const uint64_t i = t0 + gt * kItems + q;
                    if (i < N && cnt[i] < before)
                        cnt[i] = before;
                }
                before = before > cmax[gt] ? before : cmax[gt];
            }
        }
    }
    // tile_index_c: replication numbers, inclusive extra-copy prefix, zero
    // slots by rank
    uint32_t prev = 0, ex = 0, zx = 0;
    for (uint64_t i = 0; i < N; ++i) {
        rep[i] = rep_from_counts(cnt[i], prev);
        prev = cnt[i];
        ex += extra_copies(rep[i]);
        E[i] = ex;
        idx[i] = (uint32_t)i;
        if (rep[i] == 0) {
            if (zx < N)
                zpos[zx] = (uint32_t)i;
            else
                ++bad;
            ++zx;
        }
    }
    // tile_scatter_ranks
    for (uint64_t k = 0; k < nt; ++k) {
        const uint64_t t0 = k * kTile;
        const uint32_t n = (uint32_t)(N - t0 < kTile ? N - t0 : kTile);
        const uint32_t beg = t0 == 0 ? 0u : E[t0 - 1];
        const uint32_t end =
            guarded ? rank_end(E[t0 + n - 1], N) : E[t0 + n - 1];
        for (uint32_t r = beg; r < end; ++r) {
            const uint32_t s = rank_source(&E[t0], n, r);
            if (s >= n)
                ++bad;
            if (r >= N) { // zpos read past its end
                ++bad;
                break; // the rest of the range is past it too
            }
            const uint32_t z = zpos[r];
            if (guarded) {
                if (z < N)
                    idx[z] = (uint32_t)(t0 + s);
            } else if (z >= N) { // idx write past its end
                ++bad;
            } else {
                idx[z] = (uint32_t)(t0 + s);
            }
        }
    }
    return bad;
}

} // extern "C"
