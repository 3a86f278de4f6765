// The resampling kernels' index arithmetic (mckl_amd/csrc/smc_index.hpp)
// run on the CPU over a whole particle array, tile by tile, with the
// workgroup scans done sequentially: the stages of the fused samplers'
// tile_counts_l (clamp and monotone counts), tile_index_c (replication
// numbers, extra-copy prefix E, zero-slot positions zpos) and
// tile_scatter_ranks (each extra-copy rank's source, written to its zero
// slot).  Every array access the kernels make is checked against the
// array's length.  `guarded` = 0 replays the form before the count clamp and
// the rank bound (raw counts, ranks up to E of the tile's last particle),
// to show which accesses the guards keep inside the arrays.
// rs_guard_src_run replays the plain sampler's form (tile_index_local and tile_index_store, parent
// codes resolved through the rank -> source table), and guard_count_check
// pins the guarded counts of smc_index.hpp to brute force.  Test
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

// The plain sampler's form of the same transform: tile_index_local/store (zero
// slots coded by their global zero-slot rank, each tile writing the source
// of its extra-copy ranks to src at its look-back offset) and
// resolve_parent in the next move.  Counts as rs_guard_run (guarded).
// Outputs idx[N] (resolved parents), rep[N]; returns the number of
// out-of-range accesses attempted.
uint64_t rs_guard_src_run(uint64_t N, const uint32_t *raw, const uint32_t *clo,
    const uint32_t *chi, uint32_t *idx, uint32_t *rep)
{
    uint64_t bad = 0;
    const uint64_t nt = (N + kTile - 1) / kTile;
    std::vector<uint32_t> cnt(N), code(N), src(N, 0u);
    for (uint64_t k = 0; k < nt; ++k) {
        const uint64_t t0 = k * kTile;
        const uint64_t tend = t0 + kTile < N ? t0 + kTile : N;
        const uint32_t lo = k == 0 ? 0u : clo[k];
        const uint32_t hi = k + 1 == nt ? (uint32_t)N : chi[k];
        uint32_t before = lo;
        for (uint32_t gt = 0; gt < kGroup; ++gt) {
            uint32_t m = lo;
            for (uint32_t q = 0; q < kItems; ++q) {
                const uint64_t i = t0 + gt * kItems + q;
                if (i >= N)
                    continue;
                uint32_t v = i == tend - 1 ? hi : raw[i];
                v = clamp_count(v, lo, hi);
                m = m > v ? m : v;
                cnt[i] = m < before ? before : m;
            }
            before = before > m ? before : m;
        }
    }
    // tile totals, then each tile's exclusive offsets (the look-back)
    std::vector<uint32_t> te(nt, 0u), tz(nt, 0u);
    {
        uint32_t prev = 0;
        for (uint64_t i = 0; i < N; ++i) {
            if (i % kTile == 0)
                prev = i == 0 ? 0u : clo[i / kTile];
            const uint32_t r = rep_from_counts(cnt[i], prev);
            prev = cnt[i];
            te[i / kTile] += extra_copies(r);
            tz[i / kTile] += r == 0 ? 1u : 0u;
        }
    }
    uint32_t eoff = 0, zoff = 0;
    for (uint64_t k = 0; k < nt; ++k) {
        const uint64_t t0 = k * kTile;
        const uint32_t n = (uint32_t)(N - t0 < kTile ? N - t0 : kTile);
        // the fused kernel's first count of a tile follows the tile's own
        // lower boundary (tile_counts_c hands over clo)
        uint32_t prev = k == 0 ? 0u : clo[k];
        std::vector<uint32_t> Es(n);
        uint32_t ex = 0, zx = zoff;
        for (uint32_t j = 0; j < n; ++j) {
            const uint64_t i = t0 + j;
            rep[i] = rep_from_counts(cnt[i], prev);
            prev = cnt[i];
            ex += extra_copies(rep[i]);
            Es[j] = ex;
            code[i] = (uint32_t)i;
            if (rep[i] == 0)
                code[i] = 0x80000000u | zx++;
        }
        const uint32_t te_k = n ? Es[n - 1] : 0u;
        const uint32_t end = rank_end(te_k, eoff < N ? N - eoff : 0u);
        // tile_index_local: up to kTile extra copies, each source marks
        // its first rank and a max-scan over the ranks gives every rank's
        // source; beyond that, a binary search per rank in the store phase
        std::vector<uint32_t> srcl;
        if (te_k <= kTile) {
            srcl.assign(kTile, 0u);
            uint32_t e0 = 0;
            for (uint32_t j = 0; j < n; ++j) {
                if (rep[t0 + j] > 1 && e0 < kTile)
                    srcl[e0] = j;
                e0 = Es[j];
            }
            for (uint32_t r = 1; r < kTile; ++r)
                srcl[r] = srcl[r] > srcl[r - 1] ? srcl[r] : srcl[r - 1];
        }
        for (uint32_t r = 0; r < end; ++r) {
            const uint64_t g = (uint64_t)eoff + r;
            const uint32_t sj =
                te_k <= kTile ? srcl[r] : rank_source(Es.data(), n, r);
            if (sj >= n)
                ++bad;
            if (g >= N) // src write past its end
                ++bad;
            else
                src[g] = (uint32_t)(t0 + sj);
        }
        eoff += te[k];
        zoff += tz[k];
    }
    // resolve_parent
    for (uint64_t i = 0; i < N; ++i) {
        uint32_t a = code[i];
        if (a & 0x80000000u) {
            const uint32_t r = a & ~0x80000000u;
            if (r >= N) { // src read past its end
                ++bad;
                a = 0;
            } else {
                a = src[r];
            }
        }
        idx[i] = a < N ? a : (uint32_t)(N - 1);
    }
    return bad;
}

// The guarded counts of smc_index.hpp against brute force.  Uniforms of M
// (systematic: u_j = fl(fl(j delta) + u0d), u0d = fl(u01 delta); stratified:
// u_j = fl(delta fl(U_j + j)) with U_j = ustrat[j]) are counted below
// cumulative weights S placed on, next to and between uniforms (trial t
// takes Sf = a uniform moved by a few ulps, or a uniform spacing point);
// whenever a guarded count claims a result, the brute-force counts at S =
// Sf - mg, Sf, Sf + mg and at the ends of the evaluation bound must all
// equal it.  Returns the number of disagreements; *nok counts the claims.
static uint32_t count_below(const std::vector<double> &u, double S)
{
    uint32_t lo = 0, hi = (uint32_t)u.size();
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (u[mid] < S)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

uint64_t guard_count_check(int stratified, uint32_t M, double u01,
    const double *ustrat, const double *Sf_in, uint64_t ntrial,
    double margin, uint64_t *nok)
{
    const double delta = 1.0 / (double)M;
    const double u0d = u01 * delta;
    std::vector<double> u(M);
    for (uint32_t j = 0; j < M; ++j) {
        if (stratified) {
            const double a = ustrat[j] + (double)j;
            u[j] = delta * a;
        } else {
            const double a = (double)j * delta;
            u[j] = a + u0d;
        }
    }
    uint64_t bad = 0, ok = 0;
    for (uint64_t t = 0; t < ntrial; ++t) {
        const double Sf = Sf_in[t];
        const double mg = Sf * margin;
        uint32_t c = 0;
        bool claimed;
        if (stratified) {
            const int g = count_strat_guarded(Sf, mg, M, c);
            claimed = g == 1;
            if (g == 2) {
                const double uc = u[c];
                if (uc < Sf - mg) {
                    ++c;
                    claimed = true;
                } else {
                    claimed = uc >= Sf + mg;
                }
            }
        } else {
            claimed = count_sys_guarded(Sf, mg, u0d, M, c);
        }
        if (!claimed)
            continue;
        ++ok;
        const double probe[5] = {Sf - mg, Sf, Sf + mg,
            Sf * (1.0 - 17.0 * kUnitRoundoff), Sf * (1.0 + 17.0 * kUnitRoundoff)};
        for (double S : probe)
            if (count_below(u, S) != c)
                ++bad;
    }
    *nok = ok;
    return bad;
}

} // extern "C"
