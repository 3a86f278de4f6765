/*
 * mckl_oracle.c -- CPU restatement of MCKL's Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline port).
 * The product in mckl_amd/ never links this file.  Build flags are pinned in
 * oracle/Makefile: -O2 -ffp-contract=off, no -march, so every float/double
 * operation rounds once, exactly as the reference does when it is built the
 * same way (no FMA contraction, SSE2 scalar math, glibc 2.35 libm).
 *
 * Reference paths are relative to the zhouyan/MCKL source tree.
 */
#include "mckl_oracle.h"

#include <math.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== */
/* Philox4x32-10                                                          */
/* random/internal/philox_generic_4x.hpp:40-79 (round),                   */
/* random/internal/philox_constants.hpp:52-56,82-87 (constants),          */
/* random/internal/philox_common.hpp:44-73 (32x32->64 hi/lo).             */
/* The counter (array<uint64,2>) is read as four little-endian u32 words. */
/* ===================================================================== */
void mo_philox4x32_block(
    const uint32_t key[2], const uint64_t ctr[2], uint32_t out[4])
{
    uint32_t s0 = (uint32_t)ctr[0];
    uint32_t s1 = (uint32_t)(ctr[0] >> 32);
    uint32_t s2 = (uint32_t)ctr[1];
    uint32_t s3 = (uint32_t)(ctr[1] >> 32);
    uint32_t k0 = key[0];
    uint32_t k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p2 = (uint64_t)s2 * 0xCD9E8D57u;
        uint64_t p0 = (uint64_t)s0 * 0xD2511F53u;
        uint32_t hi2 = (uint32_t)(p2 >> 32), lo2 = (uint32_t)p2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t rk0 = k0 + 0x9E3779B9u * (uint32_t)r;
        uint32_t rk1 = k1 + 0xBB67AE85u * (uint32_t)r;
        uint32_t n0 = rk0 ^ s1 ^ hi2;
        uint32_t n2 = rk1 ^ s3 ^ hi0;
        s0 = n0;
        s1 = lo2;
        s2 = n2;
        s3 = lo0;
    }
    out[0] = s0;
    out[1] = s1;
    out[2] = s2;
    out[3] = s3;
}

/* ===================================================================== */
/* Threefry4x64-20                                                        */
/* random/threefry.hpp:109-121 (parity word, zero tweak),                 */
/* random/internal/threefry_common.hpp:212-247 (key injection),           */
/* random/internal/threefry_generic_4x64.hpp:40-108 (round schedule),     */
/* random/internal/threefry_constants.hpp:83-89 (rotations).              */
/* ===================================================================== */
static const int mo_tf_r0[8] = {14, 52, 23, 5, 25, 46, 58, 32};
static const int mo_tf_r1[8] = {16, 57, 40, 37, 33, 12, 22, 32};

static uint64_t mo_rotl64(uint64_t x, int l)
{
    return (x << l) | (x >> (64 - l));
}

void mo_threefry4x64_block(
    const uint64_t key[4], const uint64_t ctr[4], uint64_t out[4])
{
    uint64_t par[8];
    par[0] = key[0];
    par[1] = key[1];
    par[2] = key[2];
    par[3] = key[3];
    par[4] = 0x1BD11BDAA9FC1A22ull ^ key[0] ^ key[1] ^ key[2] ^ key[3];
    par[5] = 0; /* tweak t0 */
    par[6] = 0; /* tweak t1 */
    par[7] = 0; /* t0 ^ t1 */
    uint64_t s[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    for (int inj = 0; inj <= 5; ++inj) {
        /* injection number S = inj, applied before rounds 4*inj+1.. */
        const unsigned S = (unsigned)inj;
        s[0] += par[(S + 0) % 5];
        s[1] += par[(S + 1) % 5] + par[5 + S % 3];
        s[2] += par[(S + 2) % 5] + par[5 + (S + 1) % 3];
        s[3] += par[(S + 3) % 5] + (uint64_t)S;
        if (inj == 5)
            break;
        for (int q = 0; q < 4; ++q) {
            const int r = inj * 4 + q; /* round N-1 */
            const int l0 = mo_tf_r0[r % 8];
            const int l1 = mo_tf_r1[r % 8];
            s[0] += s[1];
            s[2] += s[3];
            uint64_t t1 = mo_rotl64(s[1], l0);
            uint64_t t3 = mo_rotl64(s[3], l1);
            s[1] = s[2] ^ t3;
            s[3] = s[0] ^ t1;
        }
    }
    out[0] = s[0];
    out[1] = s[1];
    out[2] = s[2];
    out[3] = s[3];
}

/* ===================================================================== */
/* Threefry2x32-20 and Threefry2x64-20 (the SeedGenerator's randomizers)  */
/* random/threefry.hpp:109-121 (parity word, zero tweak),                 */
/* random/internal/threefry_common.hpp:212-247 (key box: word 0 gets      */
/* par[S % 3], word 1 par[(S + 1) % 3] + S; the tweak is zero),           */
/* random/internal/threefry_generic_2x64.hpp:40-80 (round: s0 += s1;      */
/* s1 = rotl(s1, R[(N-1) % 8]) ^ s0),                                     */
/* random/internal/threefry_constants.hpp:46-80 (parity, rotations).      */
/* ===================================================================== */
static const int mo_tf2x32_r[8] = {13, 15, 26, 6, 17, 29, 16, 24};
static const int mo_tf2x64_r[8] = {16, 42, 12, 31, 16, 32, 24, 21};

static uint32_t mo_rotl32_(uint32_t x, int l)
{
    return (x << l) | (x >> (32 - l));
}

void mo_threefry2x32_block(
    const uint32_t key[2], const uint32_t in[2], uint32_t out[2])
{
    const uint32_t par[3] = {key[0], key[1], 0x1BD11BDAu ^ key[0] ^ key[1]};
    uint32_t s0 = in[0], s1 = in[1];
    for (unsigned S = 0; S <= 5; ++S) {
        s0 += par[S % 3];
        s1 += par[(S + 1) % 3] + S;
        if (S == 5)
            break;
        for (int q = 0; q < 4; ++q) {
            s0 += s1;
            s1 = mo_rotl32_(s1, mo_tf2x32_r[(S * 4 + q) % 8]);
            s1 ^= s0;
        }
    }
    out[0] = s0;
    out[1] = s1;
}

void mo_threefry2x64_block(
    const uint64_t key[2], const uint64_t in[2], uint64_t out[2])
{
    const uint64_t par[3] = {
        key[0], key[1], 0x1BD11BDAA9FC1A22ull ^ key[0] ^ key[1]};
    uint64_t s0 = in[0], s1 = in[1];
    for (unsigned S = 0; S <= 5; ++S) {
        s0 += par[S % 3];
        s1 += par[(S + 1) % 3] + S;
        if (S == 5)
            break;
        for (int q = 0; q < 4; ++q) {
            s0 += s1;
            s1 = mo_rotl64(s1, mo_tf2x64_r[(S * 4 + q) % 8]);
            s1 ^= s0;
        }
    }
    out[0] = s0;
    out[1] = s1;
}

/* ===================================================================== */
/* SeedGenerator<key_type>::get(), as used by RNGSetVector::reset()       */
/* (random/seed.hpp:54-341, random/rng_set.hpp:81-120; SeedType of a      */
/* counter engine is its key_type, random/counter.hpp:291-296).           */
/*   Philox4x32  key u32[2]: seed_type u64, one word: s %= maxs; s *= np; */
/*               s += rank; randomized by Threefry2x32 (key 0) of the     */
/*               little-endian words (s_lo, s_hi)           (:190-199)    */
/*   ARS         key u32[4]: two u64 words {s, rank}, Threefry2x64        */
/*   Threefry4x64 key u64[4]: {s, 0, 0, rank}, Threefry4x64   (:180-188)  */
/* The internal seed starts at 1 (:165) and is post-incremented per get.  */
/* Output: n keys of 4 u64 in mo_engine.key layout (u32 words zero-       */
/* extended for the 32-bit key types).                                    */
/* ===================================================================== */
void mo_seed_keys(int kind, uint64_t seed, uint64_t np, uint64_t rank,
    int randomize, size_t n, uint64_t *keys)
{
    const uint64_t zero[4] = {0, 0, 0, 0};
    uint64_t maxs = UINT64_MAX;
    if (kind == MO_PHILOX4X32 && np > 1)
        maxs = (UINT64_MAX - rank) / np + 1;
    for (size_t i = 0; i < n; ++i) {
        uint64_t s = seed + (uint64_t)i;
        uint64_t *k = keys + 4 * i;
        k[0] = k[1] = k[2] = k[3] = 0;
        if (kind == MO_PHILOX4X32) {
            s %= maxs;
            s *= np;
            s += rank;
            uint32_t w[2] = {(uint32_t)s, (uint32_t)(s >> 32)};
            if (randomize) {
                const uint32_t z[2] = {0, 0};
                mo_threefry2x32_block(z, w, w);
            }
            k[0] = w[0];
            k[1] = w[1];
        } else if (kind == MO_ARS) {
            uint64_t w[2] = {s, rank};
            if (randomize)
                mo_threefry2x64_block(zero, w, w);
            k[0] = (uint32_t)w[0];
            k[1] = (uint32_t)(w[0] >> 32);
            k[2] = (uint32_t)w[1];
            k[3] = (uint32_t)(w[1] >> 32);
        } else {
            uint64_t w[4] = {s, 0, 0, rank};
            if (randomize)
                mo_threefry4x64_block(zero, w, w);
            k[0] = w[0];
            k[1] = w[1];
            k[2] = w[2];
            k[3] = w[3];
        }
    }
}

/* ===================================================================== */
/* AES round function and ARS-5                                           */
/* random/internal/aes_generic.hpp:254-306 (ARS round keys, Weyl),        */
/* random/internal/aes_generic.hpp:413-491 (T-table round, last round),   */
/* random/internal/aes_constants.hpp:41 (T0 | T1 | T2 | T3 | S layout),   */
/* random/internal/aes_constants.hpp:289-294 (Weyl constants).            */
/* The tables are derived here from GF(2^8) arithmetic (FIPS-197 5.1.1).  */
/* ===================================================================== */
static uint8_t mo_gf_mul(uint8_t a, uint8_t b)
{
    uint8_t p = 0;
    for (int i = 0; i < 8; ++i) {
        if (b & 1)
            p ^= a;
        uint8_t hi = (uint8_t)(a & 0x80);
        a = (uint8_t)(a << 1);
        if (hi)
            a ^= 0x1B;
        b >>= 1;
    }
    return p;
}

void mo_aes_tables(uint32_t t0[256], uint8_t sbox[256])
{
    for (int x = 0; x < 256; ++x) {
        uint8_t inv = 0;
        if (x != 0) {
            for (int y = 1; y < 256; ++y) {
                if (mo_gf_mul((uint8_t)x, (uint8_t)y) == 1) {
                    inv = (uint8_t)y;
                    break;
                }
            }
        }
        uint8_t b = inv;
        uint8_t s = (uint8_t)(b ^ ((b << 1) | (b >> 7)) ^ ((b << 2) | (b >> 6)) ^
            ((b << 3) | (b >> 5)) ^ ((b << 4) | (b >> 4)) ^ 0x63);
        sbox[x] = s;
        uint32_t s2 = mo_gf_mul(s, 2), s3 = mo_gf_mul(s, 3);
        t0[x] = s2 | ((uint32_t)s << 8) | ((uint32_t)s << 16) | (s3 << 24);
    }
}

static uint32_t mo_t0[256];
static uint8_t mo_sbox[256];
static int mo_tables_ready = 0;

static void mo_ensure_tables(void)
{
    if (!mo_tables_ready) {
        mo_aes_tables(mo_t0, mo_sbox);
        mo_tables_ready = 1;
    }
}

static uint32_t mo_rotl32(uint32_t x, int l)
{
    return l == 0 ? x : (x << l) | (x >> (32 - l));
}

static void mo_aes_enc(uint32_t s[4], const uint32_t rk[4])
{
    uint32_t t[4];
    for (int j = 0; j < 4; ++j) {
        t[j] = rk[j] ^ mo_t0[s[j] & 0xFF] ^
            mo_rotl32(mo_t0[(s[(j + 1) & 3] >> 8) & 0xFF], 8) ^
            mo_rotl32(mo_t0[(s[(j + 2) & 3] >> 16) & 0xFF], 16) ^
            mo_rotl32(mo_t0[(s[(j + 3) & 3] >> 24) & 0xFF], 24);
    }
    memcpy(s, t, sizeof(t));
}

static void mo_aes_enclast(uint32_t s[4], const uint32_t rk[4])
{
    uint32_t t[4];
    for (int j = 0; j < 4; ++j) {
        t[j] = rk[j] ^ (uint32_t)mo_sbox[s[j] & 0xFF] ^
            ((uint32_t)mo_sbox[(s[(j + 1) & 3] >> 8) & 0xFF] << 8) ^
            ((uint32_t)mo_sbox[(s[(j + 2) & 3] >> 16) & 0xFF] << 16) ^
            ((uint32_t)mo_sbox[(s[(j + 3) & 3] >> 24) & 0xFF] << 24);
    }
    memcpy(s, t, sizeof(t));
}

void mo_ars_block(const uint32_t key[4], const uint64_t ctr[2], uint32_t out[4])
{
    mo_ensure_tables();
    const uint64_t K0 = (uint64_t)key[0] | ((uint64_t)key[1] << 32);
    const uint64_t K1 = (uint64_t)key[2] | ((uint64_t)key[3] << 32);
    uint32_t rk[6][4];
    for (int r = 0; r <= 5; ++r) {
        uint64_t k0 = K0 + 0x9E3779B97F4A7C15ull * (uint64_t)r;
        uint64_t k1 = K1 + 0xBB67AE8584CAA73Bull * (uint64_t)r;
        rk[r][0] = (uint32_t)k0;
        rk[r][1] = (uint32_t)(k0 >> 32);
        rk[r][2] = (uint32_t)k1;
        rk[r][3] = (uint32_t)(k1 >> 32);
    }
    uint32_t s[4] = {(uint32_t)ctr[0], (uint32_t)(ctr[0] >> 32),
        (uint32_t)ctr[1], (uint32_t)(ctr[1] >> 32)};
    for (int j = 0; j < 4; ++j)
        s[j] ^= rk[0][j];
    for (int r = 1; r < 5; ++r)
        mo_aes_enc(s, rk[r]);
    mo_aes_enclast(s, rk[5]);
    memcpy(out, s, sizeof(s));
}

/* FIPS-197 AES-128 with the same round function, used only to pin the round
 * against the FIPS-197 Appendix C vector that example/random/src/
 * random_aes.cpp:49-86 checks.  Key expansion: FIPS-197 5.2. */
void mo_aes128_encrypt(const uint8_t key[16], const uint8_t in[16], uint8_t out[16])
{
    mo_ensure_tables();
    uint32_t w[44];
    for (int i = 0; i < 4; ++i)
        w[i] = (uint32_t)key[4 * i] | ((uint32_t)key[4 * i + 1] << 8) |
            ((uint32_t)key[4 * i + 2] << 16) | ((uint32_t)key[4 * i + 3] << 24);
    uint8_t rcon = 1;
    for (int i = 4; i < 44; ++i) {
        uint32_t t = w[i - 1];
        if (i % 4 == 0) {
            t = (t >> 8) | (t << 24); /* RotWord on little-endian words */
            t = (uint32_t)mo_sbox[t & 0xFF] |
                ((uint32_t)mo_sbox[(t >> 8) & 0xFF] << 8) |
                ((uint32_t)mo_sbox[(t >> 16) & 0xFF] << 16) |
                ((uint32_t)mo_sbox[(t >> 24) & 0xFF] << 24);
            t ^= rcon;
            rcon = mo_gf_mul(rcon, 2);
        }
        w[i] = w[i - 4] ^ t;
    }
    uint32_t s[4];
    for (int i = 0; i < 4; ++i)
        s[i] = ((uint32_t)in[4 * i] | ((uint32_t)in[4 * i + 1] << 8) |
                   ((uint32_t)in[4 * i + 2] << 16) |
                   ((uint32_t)in[4 * i + 3] << 24)) ^
            w[i];
    for (int r = 1; r < 10; ++r)
        mo_aes_enc(s, w + 4 * r);
    mo_aes_enclast(s, w + 40);
    for (int i = 0; i < 4; ++i) {
        out[4 * i] = (uint8_t)s[i];
        out[4 * i + 1] = (uint8_t)(s[i] >> 8);
        out[4 * i + 2] = (uint8_t)(s[i] >> 16);
        out[4 * i + 3] = (uint8_t)(s[i] >> 24);
    }
}

/* ===================================================================== */
/* Counter engine                                                         */
/* random/counter.hpp:85-114 (seed/key), :126-165 (scalar and bulk        */
/* draws), :178-199 (discard), :282-287 (reset).                          */
/* random/internal/increment_generic.hpp:73-103 (increment by one),       */
/* :208-217 (increment by n, including its carry test).                   */
/* ===================================================================== */
static int mo_ctr_words(int kind)
{
    return kind == MO_THREEFRY4X64 ? 4 : 2;
}

static unsigned mo_block_bytes(int kind)
{
    return kind == MO_THREEFRY4X64 ? 32u : 16u;
}

unsigned mo_engine_words(const mo_engine *e)
{
    return mo_block_bytes(e->kind) / (unsigned)(e->rbits / 8);
}

/* Textual state, CounterEngine operator<< (random/counter.hpp:235-249):
 * result_ (M words), ctr_, generator_, index_, each std::array printed as
 * "size v0 v1 ..." (internal/iostream.hpp:45-63).  The generators print
 * Philox key_ (philox.hpp:145-156), Threefry par_ = key, parity, three
 * tweak words (threefry.hpp:170-181, :109-121), ARS key_ (aes_key_seq.hpp:
 * 162-173).  Returns the length written (NUL-terminated) or -1. */
int mo_engine_text(const mo_engine *e, char *text, size_t cap)
{
    char tmp[1024];
    int n = 0;
    const unsigned M = mo_engine_words(e);
    n += snprintf(tmp + n, sizeof(tmp) - n, "%u", M);
    for (unsigned i = 0; i < M; ++i) {
        uint64_t v = 0;
        memcpy(&v, e->buf + i * (e->rbits / 8), (size_t)(e->rbits / 8));
        n += snprintf(tmp + n, sizeof(tmp) - n, " %llu", (unsigned long long)v);
    }
    const int K = e->kind == MO_THREEFRY4X64 ? 4 : 2;
    n += snprintf(tmp + n, sizeof(tmp) - n, " %d", K);
    for (int i = 0; i < K; ++i)
        n += snprintf(tmp + n, sizeof(tmp) - n, " %llu",
            (unsigned long long)e->ctr[i]);
    if (e->kind == MO_PHILOX4X32) {
        n += snprintf(tmp + n, sizeof(tmp) - n, " 2 %llu %llu",
            (unsigned long long)(uint32_t)e->key[0],
            (unsigned long long)(uint32_t)e->key[1]);
    } else if (e->kind == MO_ARS) {
        n += snprintf(tmp + n, sizeof(tmp) - n, " 4");
        for (int i = 0; i < 4; ++i)
            n += snprintf(tmp + n, sizeof(tmp) - n, " %llu",
                (unsigned long long)(uint32_t)e->key[i]);
    } else {
        const uint64_t par = 0x1BD11BDAA9FC1A22ull ^ e->key[0] ^ e->key[1] ^
            e->key[2] ^ e->key[3];
        n += snprintf(tmp + n, sizeof(tmp) - n, " 8");
        for (int i = 0; i < 4; ++i)
            n += snprintf(tmp + n, sizeof(tmp) - n, " %llu",
                (unsigned long long)e->key[i]);
        n += snprintf(tmp + n, sizeof(tmp) - n, " %llu 0 0 0",
            (unsigned long long)par);
    }
    n += snprintf(tmp + n, sizeof(tmp) - n, " %u", e->index);
    if ((size_t)n + 1 > cap)
        return -1;
    memcpy(text, tmp, (size_t)n + 1);
    return n;
}

static void mo_increment1(uint64_t *ctr, int k)
{
    for (int i = 0; i < k; ++i)
        if (++ctr[i] != 0)
            return;
}

void mo_increment(uint64_t *ctr, int k, uint64_t nskip)
{
    /* increment_generic.hpp:208-217: the carry test is a strict '<'. */
    if (ctr[0] < UINT64_MAX - nskip) {
        ctr[0] += nskip;
    } else {
        ctr[0] += nskip;
        mo_increment1(ctr + 1, k - 1);
    }
}

static void mo_block(const mo_engine *e, const uint64_t *ctr, uint8_t *out)
{
    if (e->kind == MO_PHILOX4X32) {
        uint32_t key[2] = {(uint32_t)e->key[0], (uint32_t)e->key[1]};
        uint32_t r[4];
        mo_philox4x32_block(key, ctr, r);
        memcpy(out, r, 16);
    } else if (e->kind == MO_THREEFRY4X64) {
        uint64_t r[4];
        mo_threefry4x64_block(e->key, ctr, r);
        memcpy(out, r, 32);
    } else {
        uint32_t key[4] = {(uint32_t)e->key[0], (uint32_t)e->key[1],
            (uint32_t)e->key[2], (uint32_t)e->key[3]};
        uint32_t r[4];
        mo_ars_block(key, ctr, r);
        memcpy(out, r, 16);
    }
}

/* generator_(ctr_, r): increment then encrypt, counter.hpp:126-134 */
static void mo_gen_next(mo_engine *e, uint8_t *out)
{
    mo_increment1(e->ctr, mo_ctr_words(e->kind));
    mo_block(e, e->ctr, out);
}

void mo_engine_init(mo_engine *e, int kind, int rbits, const uint64_t *key)
{
    memset(e, 0, sizeof(*e));
    e->kind = kind;
    e->rbits = rbits;
    int nk = kind == MO_PHILOX4X32 ? 2 : 4;
    for (int i = 0; i < nk; ++i)
        e->key[i] = key[i];
    e->index = mo_engine_words(e);
}

void mo_engine_seed(mo_engine *e, uint64_t s)
{
    /* counter.hpp:97-103: key = {s, 0, ...} truncated to the key word type;
     * the argument itself is result_type. */
    uint64_t key[4] = {0, 0, 0, 0};
    uint64_t sv = e->rbits == 32 ? (uint32_t)s : s;
    key[0] = e->kind == MO_THREEFRY4X64 ? sv : (uint32_t)sv;
    mo_engine_init(e, e->kind, e->rbits, key);
}

void mo_engine_set_ctr(mo_engine *e, const uint64_t *ctr)
{
    for (int i = 0; i < mo_ctr_words(e->kind); ++i)
        e->ctr[i] = ctr[i];
    e->index = mo_engine_words(e);
}

static void mo_rand_bytes(mo_engine *e, size_t n, uint8_t *r)
{
    const unsigned rs = (unsigned)(e->rbits / 8);
    const unsigned M = mo_engine_words(e);
    const size_t remain = (size_t)(M - e->index);
    if (n <= remain) {
        memcpy(r, e->buf + (size_t)e->index * rs, n * rs);
        e->index += (unsigned)n;
        return;
    }
    memcpy(r, e->buf + (size_t)e->index * rs, remain * rs);
    r += remain * rs;
    n -= remain;
    e->index = M;
    const size_t m = n / M;
    for (size_t i = 0; i < m; ++i, r += (size_t)M * rs)
        mo_gen_next(e, r);
    n -= m * M;
    mo_gen_next(e, e->buf);
    memcpy(r, e->buf, n * rs);
    e->index = (unsigned)n;
}

void mo_rand_u32(mo_engine *e, size_t n, uint32_t *r)
{
    mo_rand_bytes(e, n, (uint8_t *)r);
}

void mo_rand_u64(mo_engine *e, size_t n, uint64_t *r)
{
    mo_rand_bytes(e, n, (uint8_t *)r);
}

uint32_t mo_next_u32(mo_engine *e)
{
    /* counter.hpp:134-142 */
    const unsigned M = mo_engine_words(e);
    if (e->index == M) {
        mo_gen_next(e, e->buf);
        e->index = 0;
    }
    uint32_t v;
    memcpy(&v, e->buf + (size_t)e->index * 4, 4);
    e->index++;
    return v;
}

/* scalar operator() of a 64-bit-result engine (counter.hpp:134-142) */
uint64_t mo_next_u64(mo_engine *e)
{
    const unsigned M = mo_engine_words(e);
    if (e->index == M) {
        mo_gen_next(e, e->buf);
        e->index = 0;
    }
    uint64_t v;
    memcpy(&v, e->buf + (size_t)e->index * 8, 8);
    e->index++;
    return v;
}

void mo_discard(mo_engine *e, uint64_t nskip)
{
    /* counter.hpp:178-199; rate = sizeof(result block)/sizeof(ctr) == 1 */
    if (nskip == 0)
        return;
    const unsigned M = mo_engine_words(e);
    const uint64_t remain = M - e->index;
    if (nskip <= remain) {
        e->index += (unsigned)nskip;
        return;
    }
    nskip -= remain;
    e->index = M;
    mo_increment(e->ctr, mo_ctr_words(e->kind), nskip / M);
    mo_gen_next(e, e->buf);
    e->index = (unsigned)(nskip % M);
}

/* ===================================================================== */
/* U01 conversions                                                        */
/* random/internal/u01_generic.hpp:48-164 (CC/CO/OC/OO fixed point),      */
/* :166-198 (Canonical); random/u01_distribution.hpp:95-113 (one u32 per  */
/* value for 32-bit engines), :115-129 (Canonical uses two u32/double).   */
/* ===================================================================== */
double mo_u01_u32_f64(int interval, uint32_t u)
{
    const double p32 = 0x1p-32, p33 = 0x1p-33;
    switch (interval) {
        case MO_CC:
            return ((double)(u & 1) + (double)u) * p32;
        case MO_CO:
            return (double)u * p32;
        case MO_OC:
            return (double)u * p32 + p32;
        default:
            return (double)u * p32 + p33;
    }
}

float mo_u01_u32_f32(int interval, uint32_t u)
{
    switch (interval) {
        case MO_CC: {
            uint32_t v = (uint32_t)(u << 1) >> 7;
            return (float)((v & 1) + v) * 0x1p-25f;
        }
        case MO_CO:
            return (float)(u >> 8) * 0x1p-24f;
        case MO_OC:
            return (float)(u >> 8) * 0x1p-24f + 0x1p-24f;
        default:
            return (float)(u >> 9) * 0x1p-23f + 0x1p-24f;
    }
}

double mo_u01_canonical_f64(uint32_t u0, uint32_t u1)
{
    return (double)u0 * 0x1p-64 + ((double)u1 * 0x1p-32 + 0.0);
}

float mo_u01_canonical_f32(uint32_t u)
{
    return (float)u * 0x1p-32f + 0.0f;
}

/* uniform_bits_distribution<UIntType>(rng, n, r), ubits = 16, 32, 64:
 * random/uniform_bits_distribution.hpp:60-107.  At least as wide as the
 * result word: rand(rng, n * ubits / rbits, result words) into the same
 * bytes (:72-80); narrower: chunks of BufferSize<UIntType> = 8192 /
 * sizeof(UIntType) result words, each truncated to UIntType (:82-97). */
void mo_rand_uniform_bits(mo_engine *e, int ubits, size_t n, void *r)
{
    if (ubits >= e->rbits) {
        if (e->rbits == 32)
            mo_rand_u32(e, n * (size_t)(ubits / 32), (uint32_t *)r);
        else
            mo_rand_u64(e, n, (uint64_t *)r);
        return;
    }
    const size_t K = 8192 / (size_t)(ubits / 8);
    uint8_t *out = (uint8_t *)r;
    uint64_t s64[4096];
    uint32_t s32[4096];
    for (;;) {
        /* one bulk call per chunk (the last one possibly empty) */
        const size_t k = n < K ? n : K;
        if (e->rbits == 32)
            mo_rand_u32(e, k, s32);
        else
            mo_rand_u64(e, k, s64);
        for (size_t i = 0; i < k; ++i) {
            const uint64_t w = e->rbits == 32 ? (uint64_t)s32[i] : s64[i];
            if (ubits == 16) {
                const uint16_t v = (uint16_t)w;
                memcpy(out + 2 * i, &v, 2);
            } else {
                const uint32_t v = (uint32_t)w;
                memcpy(out + 4 * i, &v, 4);
            }
        }
        out += k * (size_t)(ubits / 8);
        n -= k;
        if (k < K)
            break;
    }
}

/* The same conversions from one u64 word, the word type the reference
 * uses for every real type on a 64-bit-result engine
 * (U01UIntType, u01_distribution.hpp:105-113): W = 64 in
 * u01_generic.hpp:48-164, all exact; Canonical (:166-198) with Q = 1 is
 * one rounding of the 64-bit integer. */
double mo_u01_u64_f64(int interval, uint64_t u)
{
    switch (interval) {
        case MO_CC: {
            const uint64_t v = (u << 1) >> 10; /* P = 53, L = 1, R = 9 */
            return (double)((v & 1) + v) * 0x1p-54;
        }
        case MO_CO:
            return (double)(u >> 11) * 0x1p-53;
        case MO_OC:
            return (double)(u >> 11) * 0x1p-53 + 0x1p-53;
        default:
            return (double)(u >> 12) * 0x1p-52 + 0x1p-53;
    }
}

float mo_u01_u64_f32(int interval, uint64_t u)
{
    switch (interval) {
        case MO_CC: {
            const uint32_t v = (uint32_t)((u << 1) >> 39); /* P = 24, R = 38 */
            return (float)((v & 1) + v) * 0x1p-25f;
        }
        case MO_CO:
            return (float)(uint32_t)(u >> 40) * 0x1p-24f;
        case MO_OC:
            return (float)(uint32_t)(u >> 40) * 0x1p-24f + 0x1p-24f;
        default:
            return (float)(uint32_t)(u >> 41) * 0x1p-23f + 0x1p-24f;
    }
}

double mo_u01_canonical64_f64(uint64_t u) { return (double)u * 0x1p-64; }

float mo_u01_canonical64_f32(uint64_t u) { return (float)u * 0x1p-64f; }

/* random/internal/common.hpp:69-86 (BATCH_0 chunking by BufferSize),
 * internal/defines.hpp:62-67 (BufferSize = 8192 / sizeof(T)). */
void mo_rand_u01_f64(mo_engine *e, int interval, size_t n, double *r)
{
    const size_t K = 1024;
    if (e->rbits == 64) {
        uint64_t w[1024];
        for (;;) {
            size_t k = n < K ? n : K;
            mo_rand_u64(e, k, w);
            for (size_t i = 0; i < k; ++i)
                r[i] = interval == MO_CANONICAL ? mo_u01_canonical64_f64(w[i])
                                                : mo_u01_u64_f64(interval, w[i]);
            r += k;
            n -= k;
            if (k < K)
                return;
        }
    }
    uint32_t s[2 * 1024];
    const size_t Q = interval == MO_CANONICAL ? 2 : 1;
    for (;;) {
        size_t k = n < K ? n : K;
        mo_rand_u32(e, k * Q, s);
        for (size_t i = 0; i < k; ++i)
            r[i] = Q == 2 ? mo_u01_canonical_f64(s[2 * i], s[2 * i + 1])
                          : mo_u01_u32_f64(interval, s[i]);
        r += k;
        n -= k;
        if (k < K)
            break;
    }
}

void mo_rand_u01_f32(mo_engine *e, int interval, size_t n, float *r)
{
    const size_t K = 2048;
    if (e->rbits == 64) {
        uint64_t w[2048];
        for (;;) {
            size_t k = n < K ? n : K;
            mo_rand_u64(e, k, w);
            for (size_t i = 0; i < k; ++i)
                r[i] = interval == MO_CANONICAL ? mo_u01_canonical64_f32(w[i])
                                                : mo_u01_u64_f32(interval, w[i]);
            r += k;
            n -= k;
            if (k < K)
                return;
        }
    }
    uint32_t s[2048];
    for (;;) {
        size_t k = n < K ? n : K;
        mo_rand_u32(e, k, s);
        for (size_t i = 0; i < k; ++i)
            r[i] = interval == MO_CANONICAL ? mo_u01_canonical_f32(s[i])
                                            : mo_u01_u32_f32(interval, s[i]);
        r += k;
        n -= k;
        if (k < K)
            break;
    }
}

/* UniformReal: random/uniform_real_distribution.hpp:48-58 -- U01CO in
 * chunks of BufferSize<T>, then muladd(r, b - a, a) without FMA
 * (math/vmf.hpp:697-722): r * (b - a) rounded, then + a rounded. */
void mo_rand_uniform_real_f64(mo_engine *e, size_t n, double a, double b,
    double *r)
{
    const double d = b - a;
    mo_rand_u01_f64(e, MO_CO, n, r);
    for (size_t i = 0; i < n; ++i) {
        volatile double t = r[i] * d;
        r[i] = t + a;
    }
}

void mo_rand_uniform_real_f32(mo_engine *e, size_t n, float a, float b,
    float *r)
{
    const float d = b - a;
    mo_rand_u01_f32(e, MO_CO, n, r);
    for (size_t i = 0; i < n; ++i) {
        volatile float t = r[i] * d;
        r[i] = t + a;
    }
}

/* ===================================================================== */
/* Normal: chunked Box-Muller                                             */
/* random/normal_distribution.hpp:50-78 (chunk: first half s*sin, second  */
/* half s*cos), :82-101 (chunk loop and odd tail with two scalar U01OC    */
/* draws); math/vmf.hpp:1052-1077 (sincos), :670-722 (muladd = a*b+c      */
/* without FMA); math/constants.hpp:39-51,111 (2*pi constant).            */
/* ===================================================================== */
/* glibc's result y = f(x) against the correctly rounded value (the quad-  */
/* precision libquadmath value rounded once): 1 when they agree.  These    */
/* flags say which outputs depend only on correctly rounded libm results.  */
#define MO_CR_LOG 0
#define MO_CR_SIN 1
#define MO_CR_COS 2
#define MO_CR_LOG1P 3
static __float128 mo_quad_fn(int fn, __float128 x)
{
    switch (fn) {
    case MO_CR_LOG: return logq(x);
    case MO_CR_SIN: return sinq(x);
    case MO_CR_COS: return cosq(x);
    default: return log1pq(x);
    }
}
static int mo_cr_f32(int fn, float x, float y)
{
    return y == (float)mo_quad_fn(fn, (__float128)x);
}
static int mo_cr_f64(int fn, double x, double y)
{
    return y == (double)mo_quad_fn(fn, (__float128)x);
}

void mo_quad_eval(int fn, size_t n, const double *x, double *y)
{
    for (size_t i = 0; i < n; ++i)
        y[i] = (double)mo_quad_fn(fn, (__float128)x[i]);
}

/* x = hi + lo exactly as a quad value, minus f(x0) in quad, in units of   */
/* f(x0)'s magnitude: the relative error of a double-double table entry    */
void mo_quad_dd_relerr(int fn, size_t n, const double *x0, const double *hi,
    const double *lo, double *err)
{
    for (size_t i = 0; i < n; ++i) {
        const __float128 v = mo_quad_fn(fn, (__float128)x0[i]);
        const __float128 d = ((__float128)hi[i] + (__float128)lo[i]) - v;
        err[i] = v == 0 ? (double)fabsq(d) : (double)fabsq(d / v);
    }
}

static void mo_normal_chunk_f32(
    mo_engine *e, size_t n, float *r, float mean, float sd, uint8_t *cl)
{
    float s[1024];
    uint8_t okl[1024], oks[1024], okc[1024];
    const size_t nu = n / 2;
    float *u1 = r, *u2 = r + nu;
    mo_rand_u01_f32(e, MO_OC, n, r);
    for (size_t i = 0; i < nu; ++i)
        s[i] = logf(u1[i]);
    if (cl)
        for (size_t i = 0; i < nu; ++i)
            okl[i] = (uint8_t)mo_cr_f32(MO_CR_LOG, u1[i], s[i]);
    for (size_t i = 0; i < nu; ++i)
        s[i] = s[i] * -2.0f;
    for (size_t i = 0; i < nu; ++i)
        s[i] = sqrtf(s[i]);
    for (size_t i = 0; i < nu; ++i)
        u2[i] = 6.283185307179586476925286766559005768394f * u2[i];
    /* separate sin and cos passes, as vmf.hpp:1071-1076 does (a fused
     * sincos call may round differently) */
    for (size_t i = 0; i < nu; ++i)
        u1[i] = sinf(u2[i]);
    if (cl)
        for (size_t i = 0; i < nu; ++i)
            oks[i] = (uint8_t)mo_cr_f32(MO_CR_SIN, u2[i], u1[i]);
    for (size_t i = 0; i < nu; ++i) {
        const float th = u2[i];
        u2[i] = cosf(th);
        if (cl)
            okc[i] = (uint8_t)mo_cr_f32(MO_CR_COS, th, u2[i]);
    }
    if (cl)
        for (size_t i = 0; i < nu; ++i) {
            cl[i] = okl[i] & oks[i];
            cl[nu + i] = okl[i] & okc[i];
        }
    if (sd != 1.0f)
        for (size_t i = 0; i < nu; ++i)
            s[i] = sd * s[i];
    if (mean != 0.0f) {
        for (size_t i = 0; i < nu; ++i)
            u1[i] = s[i] * u1[i] + mean;
        for (size_t i = 0; i < nu; ++i)
            u2[i] = s[i] * u2[i] + mean;
    } else {
        for (size_t i = 0; i < nu; ++i)
            u1[i] = s[i] * u1[i];
        for (size_t i = 0; i < nu; ++i)
            u2[i] = s[i] * u2[i];
    }
}

/* clean (may be NULL): per output, 1 when every logf/sinf/cosf result it
 * depends on is correctly rounded */
void mo_rand_normal_f32_cr(mo_engine *e, size_t n, float mean, float sd,
    float *r, uint8_t *clean)
{
    const size_t k = 2048;
    const size_t m = n / k, l = n % k;
    for (size_t i = 0; i < m; ++i, r += k)
        mo_normal_chunk_f32(e, k, r, mean, sd, clean ? clean + i * k : NULL);
    mo_normal_chunk_f32(e, l, r, mean, sd, clean ? clean + m * k : NULL);
    if (n % 2 != 0) {
        /* two scalar U01OC draws: one result word each */
        float u = e->rbits == 64 ? mo_u01_u64_f32(MO_OC, mo_next_u64(e))
                                 : mo_u01_u32_f32(MO_OC, mo_next_u32(e));
        float v = e->rbits == 64 ? mo_u01_u64_f32(MO_OC, mo_next_u64(e))
                                 : mo_u01_u32_f32(MO_OC, mo_next_u32(e));
        const float lu = logf(u);
        const float th = 6.283185307179586476925286766559005768394f * v;
        const float c = cosf(th);
        r[l - 1] = mean + sd * sqrtf(-2 * lu) * c;
        if (clean)
            clean[n - 1] = (uint8_t)(mo_cr_f32(MO_CR_LOG, u, lu) &
                mo_cr_f32(MO_CR_COS, th, c));
    }
}

void mo_rand_normal_f32(mo_engine *e, size_t n, float mean, float sd, float *r)
{
    mo_rand_normal_f32_cr(e, n, mean, sd, r, NULL);
}

static void mo_normal_chunk_f64(
    mo_engine *e, size_t n, double *r, double mean, double sd, uint8_t *cl)
{
    double s[512];
    uint8_t okl[512], oks[512], okc[512];
    const size_t nu = n / 2;
    double *u1 = r, *u2 = r + nu;
    mo_rand_u01_f64(e, MO_OC, n, r);
    for (size_t i = 0; i < nu; ++i)
        s[i] = log(u1[i]);
    if (cl)
        for (size_t i = 0; i < nu; ++i)
            okl[i] = (uint8_t)mo_cr_f64(MO_CR_LOG, u1[i], s[i]);
    for (size_t i = 0; i < nu; ++i)
        s[i] = s[i] * -2.0;
    for (size_t i = 0; i < nu; ++i)
        s[i] = sqrt(s[i]);
    for (size_t i = 0; i < nu; ++i)
        u2[i] = 6.283185307179586 * u2[i];
    for (size_t i = 0; i < nu; ++i)
        u1[i] = sin(u2[i]);
    if (cl)
        for (size_t i = 0; i < nu; ++i)
            oks[i] = (uint8_t)mo_cr_f64(MO_CR_SIN, u2[i], u1[i]);
    for (size_t i = 0; i < nu; ++i) {
        const double th = u2[i];
        u2[i] = cos(th);
        if (cl)
            okc[i] = (uint8_t)mo_cr_f64(MO_CR_COS, th, u2[i]);
    }
    if (cl)
        for (size_t i = 0; i < nu; ++i) {
            cl[i] = okl[i] & oks[i];
            cl[nu + i] = okl[i] & okc[i];
        }
    if (sd != 1.0)
        for (size_t i = 0; i < nu; ++i)
            s[i] = sd * s[i];
    if (mean != 0.0) {
        for (size_t i = 0; i < nu; ++i)
            u1[i] = s[i] * u1[i] + mean;
        for (size_t i = 0; i < nu; ++i)
            u2[i] = s[i] * u2[i] + mean;
    } else {
        for (size_t i = 0; i < nu; ++i)
            u1[i] = s[i] * u1[i];
        for (size_t i = 0; i < nu; ++i)
            u2[i] = s[i] * u2[i];
    }
}

void mo_rand_normal_f64_cr(mo_engine *e, size_t n, double mean, double sd,
    double *r, uint8_t *clean)
{
    const size_t k = 1024;
    const size_t m = n / k, l = n % k;
    for (size_t i = 0; i < m; ++i, r += k)
        mo_normal_chunk_f64(e, k, r, mean, sd, clean ? clean + i * k : NULL);
    mo_normal_chunk_f64(e, l, r, mean, sd, clean ? clean + m * k : NULL);
    if (n % 2 != 0) {
        double u = e->rbits == 64 ? mo_u01_u64_f64(MO_OC, mo_next_u64(e))
                                  : mo_u01_u32_f64(MO_OC, mo_next_u32(e));
        double v = e->rbits == 64 ? mo_u01_u64_f64(MO_OC, mo_next_u64(e))
                                  : mo_u01_u32_f64(MO_OC, mo_next_u32(e));
        const double lu = log(u);
        const double th = 6.283185307179586 * v;
        const double c = cos(th);
        r[l - 1] = mean + sd * sqrt(-2 * lu) * c;
        if (clean)
            clean[n - 1] = (uint8_t)(mo_cr_f64(MO_CR_LOG, u, lu) &
                mo_cr_f64(MO_CR_COS, th, c));
    }
}

void mo_rand_normal_f64(
    mo_engine *e, size_t n, double mean, double sd, double *r)
{
    mo_rand_normal_f64_cr(e, n, mean, sd, r, NULL);
}

/* ===================================================================== */
/* Weight                                                                 */
/* core/weight.hpp:124-148 (set_log/add_log), :178-227 (max element,      */
/* chunked exp, sequential accumulate, inline ddot, scale by 1/accw,      */
/* ESS = accw^2/essw); internal/cblas.hpp:252-260 (ddot loop).            */
/* The log normaliser increment lmax + log(accw) is computed by the       */
/* reference and discarded (weight.hpp:198-227); it is returned here.     */
/* ===================================================================== */
static void mo_weight_normalize(size_t n, double *w, double *ess, double *lognc)
{
    double lmax = -INFINITY;
    for (size_t i = 0; i < n; ++i)
        if (lmax < w[i])
            lmax = w[i];
    double accw = 0, essw = 0;
    const size_t k = 1024;
    for (size_t c = 0; c < n; c += k) {
        size_t len = n - c < k ? n - c : k;
        double *p = w + c;
        for (size_t i = 0; i < len; ++i)
            p[i] = p[i] - lmax;
        for (size_t i = 0; i < len; ++i)
            p[i] = exp(p[i]);
        for (size_t i = 0; i < len; ++i)
            accw = accw + p[i];
        double d = 0;
        for (size_t i = 0; i < len; ++i)
            d += p[i] * p[i];
        essw += d;
    }
    const double inv = 1 / accw;
    for (size_t i = 0; i < n; ++i)
        w[i] = inv * w[i];
    *ess = accw * accw / essw;
    if (lognc)
        *lognc = lmax + log(accw);
}

void mo_weight_set_log(size_t n, const double *v, double *w, double *ess)
{
    memcpy(w, v, n * sizeof(double));
    mo_weight_normalize(n, w, ess, NULL);
}

void mo_weight_add_log(size_t n, const double *v, double *w, double *ess,
    double *lognc_inc)
{
    for (size_t i = 0; i < n; ++i)
        w[i] = log(w[i]);
    for (size_t i = 0; i < n; ++i)
        w[i] = v[i] + w[i];
    mo_weight_normalize(n, w, ess, lognc_inc);
}

/* ===================================================================== */
/* Resampling                                                             */
/* algorithm/resample.hpp:63-81 (sorted transform), :95-120 (stratified), */
/* :122-135 (systematic), :216-268 (rand forms: Canonical uniforms, one    */
/* per value, chunked by BufferSize<double> = 1024), :374-407 (u01 ->     */
/* replication), :416-466 (replication -> parent index), :515-576         */
/* (ResampleAlgorithm composition).                                       */
/* ===================================================================== */
void mo_u01_rand_sorted(mo_engine *e, size_t N, double *r)
{
    if (N == 0)
        return;
    const size_t k = 1024;
    double lmax = 0;
    double s[1024];
    for (size_t n0 = 0; n0 < N; n0 += k, r += k) {
        size_t n = N - n0 < k ? N - n0 : k;
        mo_rand_u01_f64(e, MO_CANONICAL, n, r);
        for (size_t j = 0; j < n; ++j)
            r[j] = log(r[j]);
        size_t m = N - n0;
        for (size_t j = 0; j < n; ++j, --m) {
            lmax += r[j] / (double)m;
            s[j] = lmax;
        }
        for (size_t j = 0; j < n; ++j)
            s[j] = exp(s[j]);
        for (size_t j = 0; j < n; ++j)
            r[j] = 1.0 - s[j];
    }
}

void mo_u01_rand_stratified(mo_engine *e, size_t N, double *r)
{
    const size_t k = 1024;
    const double delta = 1 / (double)N;
    for (size_t n0 = 0; n0 < N; n0 += k, r += k) {
        size_t n = N - n0 < k ? N - n0 : k;
        mo_rand_u01_f64(e, MO_CANONICAL, n, r);
        for (size_t j = 0; j < n; ++j)
            r[j] = r[j] + (double)(n0 + j);
        for (size_t j = 0; j < n; ++j)
            r[j] = delta * r[j];
    }
}

void mo_u01_rand_systematic(mo_engine *e, size_t N, double *r)
{
    uint32_t a = mo_next_u32(e);
    uint32_t b = mo_next_u32(e);
    double u01 = mo_u01_canonical_f64(a, b);
    if (N == 0)
        return;
    const double delta = 1 / (double)N;
    const double u = u01 * delta;
    for (size_t i = 0; i < N; ++i)
        r[i] = (double)i * delta + u;
}

void mo_resample_trans_u01_rep(
    size_t N, size_t M, const double *w, const double *u01, uint64_t *rep)
{
    if (N == 0)
        return;
    if (N == 1) {
        rep[0] = M;
        return;
    }
    memset(rep, 0, N * sizeof(uint64_t));
    if (M == 0)
        return;
    double accw = 0;
    size_t j = 0;
    for (size_t i = 0; i != N - 1; ++i) {
        accw += w[i];
        while (j != M && u01[j] < accw) {
            rep[i] += 1;
            ++j;
        }
    }
    rep[N - 1] = M - j;
}

/* The same merge with the cumulative weights correctly rounded instead of
 * summed sequentially: S_i = the exact sum of w_0..w_i rounded once (a quad
 * accumulator, 113 bits, rounded to double at each step).  This is the
 * device's definition (a double-double prefix rounded once); it differs
 * from the reference's sequential sums only at uniforms within their
 * rounding error of a boundary. */
void mo_resample_trans_u01_rep_cr(
    size_t N, size_t M, const double *w, const double *u01, uint64_t *rep)
{
    if (N == 0)
        return;
    if (N == 1) {
        rep[0] = M;
        return;
    }
    memset(rep, 0, N * sizeof(uint64_t));
    if (M == 0)
        return;
    __float128 acc = 0;
    size_t j = 0;
    for (size_t i = 0; i != N - 1; ++i) {
        acc += (__float128)w[i];
        const double S = (double)acc;
        while (j != M && u01[j] < S) {
            rep[i] += 1;
            ++j;
        }
    }
    rep[N - 1] = M - j;
}

void mo_resample_trans_rep_index(
    size_t N, size_t M, const uint64_t *rep, uint64_t *idx)
{
    if (N == 0 || M == 0)
        return;
    const size_t K = N < M ? N : M;
    uint64_t time = 0;
    size_t src = 0;
    for (size_t dst = 0; dst < M; ++dst) {
        if (dst < K && rep[dst] > 0) {
            idx[dst] = dst;
            continue;
        }
        /* seek, resample.hpp:432-446 */
        if (src < K && rep[src] < time + 2) {
            time = 0;
            do {
                ++src;
            } while (src < K && rep[src] < time + 2);
        }
        if (src >= K && rep[src] < time + 1) {
            time = 0;
            do {
                ++src;
            } while (src < N && rep[src] < time + 1);
        }
        idx[dst] = src;
        ++time;
    }
}

/* resample_trans_residual (algorithm/resample.hpp:333-364): integral parts
 * and normalised residuals of M W_i; returns R = M - sum of integral parts */
size_t mo_resample_trans_residual(size_t N, size_t M, const double *w,
    double *resid, uint64_t *integ)
{
    double sum_resid = 0;
    uint64_t sum_integ = 0;
    const double coeff = (double)M;
    for (size_t i = 0; i != N; ++i) {
        const double x = coeff * w[i];
        double ip;
        resid[i] = modf(x, &ip);
        integ[i] = (uint64_t)ip;
        sum_resid += resid[i];
        sum_integ += integ[i];
    }
    const double mul = 1 / sum_resid;
    for (size_t i = 0; i != N; ++i)
        resid[i] *= mul;
    return M - (size_t)sum_integ;
}

static void mo_u01_sequence(int scheme, mo_engine *e, size_t n, double *u)
{
    if (scheme == MO_MULTINOMIAL || scheme == MO_RESIDUAL)
        mo_u01_rand_sorted(e, n, u);
    else if (scheme == MO_STRATIFIED || scheme == MO_RESIDUAL_STRATIFIED)
        mo_u01_rand_stratified(e, n, u);
    else
        mo_u01_rand_systematic(e, n, u);
}

/* ResampleAlgorithm::eval (algorithm/resample.hpp:538-576): the plain
 * schemes draw M uniforms; the residual ones keep the integral parts and
 * draw R uniforms over the normalised residuals (schemes :581-593). */
void mo_resample(int scheme, size_t N, size_t M, mo_engine *e,
    const double *w, uint64_t *rep, uint64_t *idx)
{
    if (scheme >= MO_RESIDUAL) {
        double *resid = (double *)malloc((N ? N : 1) * sizeof(double));
        uint64_t *integ = (uint64_t *)malloc((N ? N : 1) * sizeof(uint64_t));
        const size_t R = mo_resample_trans_residual(N, M, w, resid, integ);
        double *u = (double *)malloc((R ? R : 1) * sizeof(double));
        mo_u01_sequence(scheme, e, R, u);
        mo_resample_trans_u01_rep(N, R, resid, u, rep);
        for (size_t i = 0; i != N; ++i)
            rep[i] += integ[i];
        free(u);
        free(integ);
        free(resid);
    } else {
        double *u = (double *)malloc((M ? M : 1) * sizeof(double));
        mo_u01_sequence(scheme, e, M, u);
        mo_resample_trans_u01_rep(N, M, w, u, rep);
        free(u);
    }
    if (idx)
        mo_resample_trans_rep_index(N, M, rep, idx);
}

/* ===================================================================== */
/* This project's SMC particle filter, restated on the CPU.               */
/*                                                                        */
/* Loop shape follows SMCSampler::do_iterate (algorithm/smc.hpp:389-406): */
/* move+weight, ESS test against threshold*N (:397), resample with the    */
/* ResampleAlgorithm and rep->index transform restated above.  The        */
/* random-number layout, the FP32 arithmetic and the log-likelihood       */
/* accounting are this project's own design (the reference has no         */
/* log-likelihood API, core/weight.hpp:198-227); the device kernels       */
/* implement the same definitions:                                        */
/*   eps(i,t): Philox4x32 block at counter {i/2, t}, key {seed lo, hi};   */
/*     u1 = U01OC(w0), u2 = U01OC(w1); r = sqrt(-2 log_f(u1));            */
/*     eps = r cos(2 pi u2) for even i, r sin(2 pi u2) for odd i.         */
/*   resampling uniforms of step t: Philox4x32 blocks at counter          */
/*     {2^63 + j/2, t}, Canonical doubles from words (0,1) / (2,3).       */
/*   log_f, exp_f, sincos2pi_f: float polynomials with explicit fused     */
/*     multiply-adds in a fixed order (identical bits on CPU and GPU).    */
/*   logZ += m_t + log S_t - log C_t, C_t = N after a resampling step     */
/*     (and at t = 1), else S_{t-1}; S_t = sum exp_f(lw_i - m_t).         */
/* ===================================================================== */
static float mo_asf(uint32_t b)
{
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static uint32_t mo_asu(float f)
{
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
}

/* log(x) for positive normal x: x = 2^e m, m in [sqrt(1/2), sqrt(2)),
 * log(1+f) = f + f^2 P(f), P degree 8 (least squares on Chebyshev nodes). */
float mo_log_f(float x)
{
    uint32_t b = mo_asu(x);
    int e = (int)(b >> 23) - 127;
    uint32_t mb = (b & 0x7FFFFFu) | 0x3F800000u;
    if (mb > 0x3FB504F3u) {
        mb -= 0x00800000u;
        e += 1;
    }
    const float f = mo_asf(mb) - 1.0f;
    float p = -0x1.4237fcp-4f;
    p = fmaf(p, f, 0x1.0696e4p-3f);
    p = fmaf(p, f, -0x1.0c524cp-3f);
    p = fmaf(p, f, 0x1.22973ap-3f);
    p = fmaf(p, f, -0x1.548882p-3f);
    p = fmaf(p, f, 0x1.99a3ecp-3f);
    p = fmaf(p, f, -0x1.000206p-2f);
    p = fmaf(p, f, 0x1.55554ep-2f);
    p = fmaf(p, f, -0x1.fffffep-2f);
    const float f2 = f * f;
    const float r = fmaf(f2, p, f);
    const float ef = (float)e;
    return fmaf(ef, 0x1.62e430p-1f, fmaf(ef, -0x1.05c610p-29f, r));
}

/* exp(x): k = rint(x log2 e), r = x - k ln2 (two fmas),
 * exp(r) = 1 + r + r^2 Q(r), Q degree 5; underflow to 0 below -87. */
float mo_exp_f(float x)
{
    if (x < -87.0f)
        return 0.0f;
    if (x > 88.0f)
        return INFINITY;
    const float k = rintf(x * 0x1.715476p+0f);
    float r = fmaf(-k, 0x1.62e430p-1f, x);
    r = fmaf(-k, -0x1.05c610p-29f, r);
    float q = 0x1.a1516cp-13f;
    q = fmaf(q, r, 0x1.6d431ap-10f);
    q = fmaf(q, r, 0x1.1110c6p-7f);
    q = fmaf(q, r, 0x1.5554eap-5f);
    q = fmaf(q, r, 0x1.555556p-3f);
    q = fmaf(q, r, 0x1.000000p-1f);
    const float p = 1.0f + fmaf(r * r, q, r);
    return p * mo_asf((uint32_t)((int)k + 127) << 23);
}

/* sin(2 pi u), cos(2 pi u) for u in [0, 1]: exact quadrant reduction
 * u = q/4 + r, |r| <= 1/8, then odd/even polynomials in phi = 2 pi r. */
void mo_sincos2pi_f(float u, float *s, float *c)
{
    const float q = rintf(u * 4.0f);
    const float r = u - q * 0.25f;
    const float ph = r * 0x1.921fb6p+2f;
    const float p2 = ph * ph;
    float sp = 0x1.6dbd24p-19f;
    sp = fmaf(sp, p2, -0x1.a013a4p-13f);
    sp = fmaf(sp, p2, 0x1.11110ep-7f);
    sp = fmaf(sp, p2, -0x1.555556p-3f);
    const float sn = fmaf(ph * p2, sp, ph);
    float cp = -0x1.2469bap-22f;
    cp = fmaf(cp, p2, 0x1.a0125ap-16f);
    cp = fmaf(cp, p2, -0x1.6c16bap-10f);
    cp = fmaf(cp, p2, 0x1.555556p-5f);
    cp = fmaf(cp, p2, -0x1.000000p-1f);
    const float cs = fmaf(p2, cp, 1.0f);
    switch ((int)q & 3) {
        case 0:
            *s = sn;
            *c = cs;
            break;
        case 1:
            *s = cs;
            *c = -sn;
            break;
        case 2:
            *s = -sn;
            *c = -cs;
            break;
        default:
            *s = -cs;
            *c = sn;
            break;
    }
}

float mo_smc_normal(uint64_t seed, uint64_t i, uint64_t t)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint64_t ctr[2] = {i >> 1, t};
    uint32_t w[4];
    mo_philox4x32_block(key, ctr, w);
    const float u1 = mo_u01_u32_f32(MO_OC, w[0]);
    const float u2 = mo_u01_u32_f32(MO_OC, w[1]);
    const float rad = sqrtf(-2.0f * mo_log_f(u1));
    float s, c;
    mo_sincos2pi_f(u2, &s, &c);
    return (i & 1) ? rad * s : rad * c;
}

double mo_smc_uniform(uint64_t seed, uint64_t t, uint64_t j)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint64_t ctr[2] = {0x8000000000000000ull + j / 2, t};
    uint32_t w[4];
    mo_philox4x32_block(key, ctr, w);
    return (j & 1) ? mo_u01_canonical_f64(w[2], w[3])
                   : mo_u01_canonical_f64(w[0], w[1]);
}

typedef struct {
    float a, s, sd0, inv_sy, cst;
} mo_model;

static void mo_model_init(int model, const double *th, mo_model *m)
{
    const double l2pi = log(2.0 * 3.141592653589793);
    if (model == MO_MODEL_CV2) {
        /* theta = (sp, sv, sy): position and velocity noise scales, the
         * observation scale; both components start N(0, 1) */
        m->s = (float)th[0];
        m->a = (float)th[1];
        m->sd0 = 1.0f;
        m->inv_sy = (float)(1.0 / th[2]);
        m->cst = (float)(-log(th[2]) - 0.5 * l2pi);
        return;
    }
    if (model == MO_MODEL_GROWTH) {
        m->a = 0.0f;
        m->s = (float)th[0];
        m->inv_sy = (float)(1.0 / th[1]);
        m->cst = (float)(-log(th[1]) - 0.5 * l2pi);
        m->sd0 = (float)th[2];
        return;
    }
    m->a = (float)th[0];
    m->s = (float)th[1];
    m->cst = (float)(-log(th[2]) - 0.5 * l2pi);
    if (model == MO_MODEL_LG) {
        m->sd0 = 1.0f;
        m->inv_sy = (float)(1.0 / th[2]);
    } else {
        m->sd0 = (float)sqrt(th[1] * th[1] / (1.0 - th[0] * th[0]));
        m->inv_sy = 0.0f;
    }
}

/* per-step observation constant: LG (float)y, SV (float)(y^2 / beta^2) */
float mo_smc_obs(int model, const double *th, double y)
{
    if (model == MO_MODEL_GROWTH || model == MO_MODEL_CV2)
        return (float)y;
    return model == MO_MODEL_LG ? (float)y : (float)(y * y / (th[2] * th[2]));
}

float mo_smc_logg(int model, const double *th, float obs, float x)
{
    mo_model m;
    mo_model_init(model, th, &m);
    if (model == MO_MODEL_GROWTH) {
        const float h = (x * x) * 0.05f;
        const float d = obs - h;
        const float z = d * m.inv_sy;
        return -0.5f * (z * z) + m.cst;
    }
    if (model == MO_MODEL_LG || model == MO_MODEL_CV2) {
        /* CV2: x is the position component */
        const float d = obs - x;
        const float z = d * m.inv_sy;
        return -0.5f * (z * z) + m.cst;
    }
    const float q = obs * mo_exp_f(-x);
    return ((-0.5f * q) - (0.5f * x)) + m.cst;
}

/* (hi, lo) += d in double-double: two-sum of the high parts, then the
 * low parts, renormalised (the device's dd_add with a zero low part) */
static void mo_dd_add_d(double *hi, double *lo, double d)
{
    const double s = *hi + d;
    const double bb = s - *hi;
    const double err = (*hi - (s - bb)) + (d - bb);
    const double l = err + *lo;
    const double h = s + l;
    *lo = l - (h - s);
    *hi = h;
}

/* particles per tile of the device samplers: the weights are scaled per
 * tile (see below), so the oracle uses the same tiling */
#define MO_SMC_TILE 2048

/* kmom > 0: the selection estimator of x^1..x^kmom after each step's move
 * and weighting (SMCEstimator at step 0, algorithm/smc.hpp:89-118,
 * :389-394): est[(t-1) kmom + j-1] = sum_i W_i x_i^j, W_i = e_i * (1/S) as
 * the resampling weights below, summed in order (the reference's dgemv). */
/* a request to record the resampling inputs and outputs of one step (the
 * first resampling step at or after t, non-residual schemes) and stop */
typedef struct {
    size_t t;      /* in: earliest step; out: the step recorded (0: none) */
    double *u01;   /* the step's uniform draws (systematic: one) */
    double *W;     /* normalised weights e_i (1/S) */
    double *u;     /* the uniform sequence the counts use */
    uint64_t *rep; /* replication numbers */
    int done;
} mo_smc_dump;

static double smc_run_impl(int model, const double *theta, const double *y,
    size_t T, size_t N, uint64_t seed, double threshold, int scheme,
    int nthreads, double *ess_hist, int *resampled_hist, int kmom, double *est,
    mo_smc_dump *dump)
{
    mo_model m;
    mo_model_init(model, theta, &m);
    /* CV2 keeps (position, velocity): component d of particle i at
     * x[d N + i], its noise the stream's normal 2 i + d */
    const size_t D = model == MO_MODEL_CV2 ? 2 : 1;
    float *x = (float *)malloc(D * N * sizeof(float));
    float *xn = (float *)malloc(D * N * sizeof(float));
    float *lw = (float *)malloc(N * sizeof(float));
    double *e = (double *)malloc(N * sizeof(double));
    double *resid = (double *)malloc(N * sizeof(double));
    float *tmax = (float *)malloc(((N + MO_SMC_TILE - 1) / MO_SMC_TILE + 1) *
        sizeof(float));
    double *u = (double *)malloc(N * sizeof(double));
    uint64_t *rep = (uint64_t *)malloc(N * sizeof(uint64_t));
    uint64_t *idx = (uint64_t *)malloc(N * sizeof(uint64_t));
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const long long Ns = (long long)N;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < Ns; ++i) {
        if (D == 2) {
            x[i] = mo_smc_normal(seed, 2 * (uint64_t)i, 0);
            x[N + i] = mo_smc_normal(seed, 2 * (uint64_t)i + 1, 0);
        } else {
            x[i] = m.sd0 * mo_smc_normal(seed, (uint64_t)i, 0);
        }
        lw[i] = 0.0f;
    }
    /* t = 1 carries lw = 0 with m_0 = 0, so C_1 = S_0 = N */
    float mprev = 0.0f;
    int resampled = 0;
    double Sprev = (double)N;
    double logz = 0;
    for (size_t t = 1; t <= T; ++t) {
        const float obs = mo_smc_obs(model, theta, y[t - 1]);
        const float drift = model == MO_MODEL_GROWTH ? (float)y[T + t - 1]
                                                     : 0.0f;
        const double C = resampled ? (double)N : Sprev;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < Ns; ++i) {
            float xp, carry;
            const size_t src = resampled ? (size_t)idx[i] : (size_t)i;
            if (resampled) {
                xp = x[idx[i]];
                carry = 0.0f;
            } else {
                xp = x[i];
                carry = lw[i] - mprev;
            }
            if (D == 2) {
                /* position += velocity + sp e0, velocity += sv e1 */
                const float e0 = mo_smc_normal(seed, 2 * (uint64_t)i, t);
                const float e1 = mo_smc_normal(seed, 2 * (uint64_t)i + 1, t);
                const float vp = x[N + src];
                const float x0 = (xp + vp) + m.s * e0;
                xn[i] = x0;
                xn[N + i] = vp + m.a * e1;
                lw[i] = carry + mo_smc_logg(model, theta, obs, x0);
                continue;
            }
            const float eps = mo_smc_normal(seed, (uint64_t)i, t);
            float xv;
            if (model == MO_MODEL_GROWTH) {
                const float q = xp / (1.0f + xp * xp);
                xv = ((0.5f * xp + 25.0f * q) + drift) + m.s * eps;
            } else {
                xv = m.a * xp + m.s * eps;
            }
            xn[i] = xv;
            lw[i] = carry + mo_smc_logg(model, theta, obs, xv);
        }
        float *tmp = x;
        x = xn;
        xn = tmp;
        /* weights e_i = a_i r_k for particle i of tile k (2048 particles):
         * a_i = exp_f(lw_i - m_k), r_k = exp_f(m_k - m), m_k the tile max of
         * lw and m the max over tiles; the product of the two floats is
         * exact in double */
        const size_t ntile = (N + MO_SMC_TILE - 1) / MO_SMC_TILE;
        float mx = -INFINITY;
        for (size_t k = 0; k < ntile; ++k) {
            float mk = -INFINITY;
            const size_t b = k * MO_SMC_TILE;
            const size_t en = b + MO_SMC_TILE < N ? b + MO_SMC_TILE : N;
            for (size_t i = b; i < en; ++i)
                if (mk < lw[i])
                    mk = lw[i];
            tmax[k] = mk;
            if (mx < mk)
                mx = mk;
        }
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < Ns; ++i) {
            const float mk = tmax[(size_t)i / MO_SMC_TILE];
            e[i] = (double)mo_exp_f(lw[i] - mk) * (double)mo_exp_f(mk - mx);
        }
        /* S: the total rounded once (the device sums in double-double);
         * S2 in plain double */
        double S = 0, Slo = 0, S2 = 0;
        for (size_t i = 0; i < N; ++i) {
            const double d = e[i];
            mo_dd_add_d(&S, &Slo, d);
            S2 += d * d;
        }
        logz += ((double)mx + log(S)) - log(C);
        if (kmom > 0) {
            const double inv = 1.0 / S;
            for (int j = 1; j <= kmom; ++j) {
                double acc = 0;
                for (size_t i = 0; i < N; ++i) {
                    double p = 1.0;
                    for (int q = 0; q < j; ++q)
                        p = p * (double)x[i];
                    acc += (e[i] * inv) * p;
                }
                est[(t - 1) * (size_t)kmom + (size_t)(j - 1)] = acc;
            }
        }
        const double ess = S * S / S2;
        if (ess_hist)
            ess_hist[t - 1] = ess;
        mprev = mx;
        Sprev = S;
        resampled = ess < threshold * (double)N;
        if (resampled_hist)
            resampled_hist[t - 1] = resampled;
        if (resampled && scheme >= MO_RESIDUAL) {
            /* residual schemes (resample.hpp:333-364, 515-593):
             * W_i = e_i (1/S), N W_i = I_i + r_i (modf), R = N -
             * sum I; the R uniforms of the base scheme against the
             * residuals' prefix (rounded once) over their total (rounded
             * once), I_i copies kept.  rep holds I until the counts. */
            const double inv = 1.0 / S;
            uint64_t Isum = 0;
            double th = 0.0, tl = 0.0;
            for (size_t i = 0; i < N; ++i) {
                const double w = (double)N * (e[i] * inv);
                const double fw = floor(w);
                rep[i] = (uint64_t)fw;
                resid[i] = w - fw;
                Isum += rep[i];
                mo_dd_add_d(&th, &tl, resid[i]);
            }
            const uint64_t R = Isum <= N ? N - Isum : 0;
            const double dR = R ? 1.0 / (double)R : 0.0;
            if (scheme == MO_RESIDUAL_SYSTEMATIC) {
                const double uu = mo_smc_uniform(seed, t, 0) * dR;
                for (size_t j = 0; j < R; ++j)
                    u[j] = (double)j * dR + uu;
            } else if (scheme == MO_RESIDUAL) {
                /* R sorted uniforms, as the multinomial step draws N */
                double lm = 0;
                for (size_t j = 0; j < R; ++j) {
                    lm += log(mo_smc_uniform(seed, t, j)) / (double)(R - j);
                    u[j] = 1.0 - exp(lm);
                }
            } else {
                for (size_t j = 0; j < R; ++j)
                    u[j] = dR * (mo_smc_uniform(seed, t, j) + (double)j);
            }
            if (dump && t >= dump->t) {
                if (scheme == MO_RESIDUAL_SYSTEMATIC)
                    dump->u01[0] = mo_smc_uniform(seed, t, 0);
                else
                    for (size_t j = 0; j < R; ++j)
                        dump->u01[j] = mo_smc_uniform(seed, t, j);
                for (size_t j = 0; j < N; ++j)
                    dump->W[j] = e[j] * inv;
                for (size_t j = 0; j < R; ++j)
                    dump->u[j] = u[j];
            }
            const double invR = R ? 1.0 / th : 0.0;
            double hi = 0.0, lo = 0.0;
            size_t j = 0;
            uint64_t prevc = 0;
            for (size_t i = 0; i < N; ++i) {
                mo_dd_add_d(&hi, &lo, resid[i]);
                uint64_t c = R;
                if (i + 1 < N) {
                    c = 0;
                    if (R) {
                        const double Si = hi * invR;
                        while (j < R && u[j] < Si)
                            ++j;
                        c = j;
                    }
                }
                rep[i] += c - prevc;
                prevc = c;
            }
            if (dump && t >= dump->t) {
                memcpy(dump->rep, rep, N * sizeof(uint64_t));
                dump->t = t;
                dump->done = 1;
                break; /* nothing after the recorded step is needed */
            }
            mo_resample_trans_rep_index(N, N, rep, idx);
        } else if (resampled) {
            const double inv = 1.0 / S;
            const double delta = 1.0 / (double)N;
            if (scheme == MO_SYSTEMATIC) {
                const double uu = mo_smc_uniform(seed, t, 0) * delta;
                for (size_t j = 0; j < N; ++j)
                    u[j] = (double)j * delta + uu;
            } else if (scheme == MO_STRATIFIED) {
                for (size_t j = 0; j < N; ++j)
                    u[j] = delta * (mo_smc_uniform(seed, t, j) + (double)j);
            } else {
                double lm = 0;
                for (size_t j = 0; j < N; ++j) {
                    lm += log(mo_smc_uniform(seed, t, j)) / (double)(N - j);
                    u[j] = 1.0 - exp(lm);
                }
            }
            if (dump && t >= dump->t) {
                if (scheme == MO_SYSTEMATIC)
                    dump->u01[0] = mo_smc_uniform(seed, t, 0);
                else
                    for (size_t j = 0; j < N; ++j)
                        dump->u01[j] = mo_smc_uniform(seed, t, j);
                for (size_t j = 0; j < N; ++j) {
                    dump->W[j] = e[j] * inv;
                    dump->u[j] = u[j];
                }
            }
            /* counts from the cumulative weights rounded once: the device
             * sums the weights in double-double (within tiles, then over
             * tile offsets), so its cumulative weight of particle i is the
             * double nearest the exact prefix; a double-double running sum
             * gives the same double here (up to ties at 2^-104).
             * cnt_i = #{u_j < fl(P_i * (1/S))}, the last particle takes N. */
            {
                double hi = 0.0, lo = 0.0;
                size_t j = 0;
                uint64_t prevc = 0;
                for (size_t i = 0; i < N; ++i) {
                    mo_dd_add_d(&hi, &lo, e[i]);
                    uint64_t c = N;
                    if (i + 1 < N) {
                        const double Si = hi * inv;
                        while (j < N && u[j] < Si)
                            ++j;
                        c = j;
                    }
                    rep[i] = c - prevc;
                    prevc = c;
                }
            }
            if (dump && t >= dump->t) {
                memcpy(dump->rep, rep, N * sizeof(uint64_t));
                dump->t = t;
                dump->done = 1;
                break; /* nothing after the recorded step is needed */
            }
            mo_resample_trans_rep_index(N, N, rep, idx);
        }
    }
    free(x);
    free(xn);
    free(lw);
    free(e);
    free(resid);
    free(tmax);
    free(u);
    free(rep);
    free(idx);
    return logz;
}

double mo_smc_run_est(int model, const double *theta, const double *y,
    size_t T, size_t N, uint64_t seed, double threshold, int scheme,
    int nthreads, double *ess_hist, int *resampled_hist, int kmom, double *est)
{
    return smc_run_impl(model, theta, y, T, N, seed, threshold, scheme,
        nthreads, ess_hist, resampled_hist, kmom, est, NULL);
}

size_t mo_smc_resample_dump(int model, const double *theta, const double *y,
    size_t T, size_t N, uint64_t seed, double threshold, int scheme,
    size_t t_first, double *u01, double *W, double *u, uint64_t *rep)
{
    if (scheme < MO_MULTINOMIAL || scheme > MO_RESIDUAL_SYSTEMATIC)
        return 0;
    mo_smc_dump d = {t_first ? t_first : 1, u01, W, u, rep, 0};
    (void)smc_run_impl(model, theta, y, T, N, seed, threshold, scheme, 1,
        NULL, NULL, 0, NULL, &d);
    return d.done ? d.t : 0;
}

double mo_smc_run(int model, const double *theta, const double *y, size_t T,
    size_t N, uint64_t seed, double threshold, int scheme, int nthreads,
    double *ess_hist, int *resampled_hist)
{
    return mo_smc_run_est(model, theta, y, T, N, seed, threshold, scheme,
        nthreads, ess_hist, resampled_hist, 0, NULL);
}
