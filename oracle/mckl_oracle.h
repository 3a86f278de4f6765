/*
 * mckl_oracle.h -- CPU restatement of MCKL's Monte Carlo hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in mckl_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Paths are relative to the zhouyan/MCKL source tree.
 *
 * Parity pinning (what checks each part, see tests/test_oracle_*.py):
 *   seed keys      reference KAT files random_seed_{2x32,1x64,4x32,2x64,
 *                  4x64,8x32}.data (plain and randomized columns).
 *   engines        reference KAT files random_{philox4x32,threefry4x64,ars}
 *                  [_64].data; FIPS-197 Appendix C for the AES round; the
 *                  compiled reference (oracle/_ref) for bulk/partial/discard.
 *   u01            closed forms (exact), edge words 0 and 2^32-1, oracle/_ref.
 *   normal         oracle/_ref (bit-exact, glibc libm on both sides);
 *                  moments against theory.
 *   weight         oracle/_ref (bit-exact); closed form on tiny inputs.
 *   resampling     oracle/_ref (bit-exact); reference invariants; brute force
 *                  on all small replication vectors.
 *   smc / kalman   the SMC restatement is this project's own device algorithm
 *                  (no reference counterpart); pinned by the Kalman filter
 *                  (closed form) within Monte Carlo error, and the Kalman
 *                  filter is pinned by brute-force Gaussian integration.
 */
#ifndef MCKL_ORACLE_H
#define MCKL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MO_PHILOX4X32 = 1,
    MO_THREEFRY4X64 = 2,
    MO_ARS = 3
};

enum { MO_CC = 0, MO_CO = 1, MO_OC = 2, MO_OO = 3, MO_CANONICAL = 4 };

/* ResampleScheme numbering, algorithm/resample.hpp:43-50 */
enum {
    MO_MULTINOMIAL = 0,
    MO_STRATIFIED = 1,
    MO_SYSTEMATIC = 2,
    MO_RESIDUAL = 3,
    MO_RESIDUAL_STRATIFIED = 4,
    MO_RESIDUAL_SYSTEMATIC = 5
};

/* Counter engine state: restates CounterEngine (random/counter.hpp:61-288).
 * buf holds the buffered block as raw bytes, index counts consumed result
 * words (M == empty). */
typedef struct {
    int kind;
    int rbits;
    uint64_t ctr[4];
    uint64_t key[4];
    uint8_t buf[32];
    unsigned index;
} mo_engine;

/* ---- generators: one block for an explicit counter value ---- */
void mo_philox4x32_block(
    const uint32_t key[2], const uint64_t ctr[2], uint32_t out[4]);
void mo_threefry4x64_block(
    const uint64_t key[4], const uint64_t ctr[4], uint64_t out[4]);
void mo_ars_block(const uint32_t key[4], const uint64_t ctr[2], uint32_t out[4]);
void mo_threefry2x32_block(
    const uint32_t key[2], const uint32_t in[2], uint32_t out[2]);
void mo_threefry2x64_block(
    const uint64_t key[2], const uint64_t in[2], uint64_t out[2]);

/* ---- SeedGenerator keys for RNGSetVector (n keys, 4 u64 each) ---- */
void mo_seed_keys(int kind, uint64_t seed, uint64_t np, uint64_t rank,
    int randomize, size_t n, uint64_t *keys);
void mo_aes_tables(uint32_t t0[256], uint8_t sbox[256]);
void mo_aes128_encrypt(const uint8_t key[16], const uint8_t in[16], uint8_t out[16]);

/* ---- engine ---- */
unsigned mo_engine_words(const mo_engine *e);
void mo_engine_init(mo_engine *e, int kind, int rbits, const uint64_t *key);
void mo_engine_seed(mo_engine *e, uint64_t s);
int mo_engine_text(const mo_engine *e, char *text, size_t cap);
void mo_engine_set_ctr(mo_engine *e, const uint64_t *ctr);
void mo_rand_u32(mo_engine *e, size_t n, uint32_t *r);
void mo_rand_u64(mo_engine *e, size_t n, uint64_t *r);
uint32_t mo_next_u32(mo_engine *e);
void mo_discard(mo_engine *e, uint64_t nskip);
void mo_increment(uint64_t *ctr, int k, uint64_t nskip);

/* ---- u01 conversions ---- */
double mo_u01_u32_f64(int interval, uint32_t u);
float mo_u01_u32_f32(int interval, uint32_t u);
double mo_u01_canonical_f64(uint32_t u0, uint32_t u1);
float mo_u01_canonical_f32(uint32_t u);
/* uniform_bits_distribution<UIntType>, ubits = 16, 32 or 64 */
void mo_rand_uniform_bits(mo_engine *e, int ubits, size_t n, void *r);
/* one u64 word per value (64-bit-result engines) */
double mo_u01_u64_f64(int interval, uint64_t u);
float mo_u01_u64_f32(int interval, uint64_t u);
double mo_u01_canonical64_f64(uint64_t u);
float mo_u01_canonical64_f32(uint64_t u);
uint64_t mo_next_u64(mo_engine *e);
void mo_rand_u01_f64(mo_engine *e, int interval, size_t n, double *r);
void mo_rand_u01_f32(mo_engine *e, int interval, size_t n, float *r);
void mo_rand_uniform_real_f64(mo_engine *e, size_t n, double a, double b,
    double *r);
void mo_rand_uniform_real_f32(mo_engine *e, size_t n, float a, float b,
    float *r);

/* ---- normal ---- */
void mo_rand_normal_f32(mo_engine *e, size_t n, float mean, float sd, float *r);
void mo_rand_normal_f64(
    mo_engine *e, size_t n, double mean, double sd, double *r);
/* the same fills; clean[i] = 1 when every libm result output i depends on
 * (log, and sin or cos) equals the correctly rounded value */
void mo_rand_normal_f32_cr(mo_engine *e, size_t n, float mean, float sd,
    float *r, uint8_t *clean);
void mo_rand_normal_f64_cr(mo_engine *e, size_t n, double mean, double sd,
    double *r, uint8_t *clean);
/* quad-precision log (fn 0), sin (1), cos (2) rounded once to double */
void mo_quad_eval(int fn, size_t n, const double *x, double *y);
/* |(hi + lo) - f(x0)| / |f(x0)| in quad precision */
void mo_quad_dd_relerr(int fn, size_t n, const double *x0, const double *hi,
    const double *lo, double *err);

/* ---- weight ---- */
void mo_weight_set_log(size_t n, const double *v, double *w, double *ess);
void mo_weight_add_log(size_t n, const double *v, double *w, double *ess,
    double *lognc_inc);

/* ---- resampling ---- */
void mo_u01_rand_sorted(mo_engine *e, size_t n, double *r);
void mo_u01_rand_stratified(mo_engine *e, size_t n, double *r);
void mo_u01_rand_systematic(mo_engine *e, size_t n, double *r);
void mo_resample_trans_u01_rep(size_t n, size_t m, const double *w,
    const double *u01, uint64_t *rep);
/* cumulative weights correctly rounded (quad accumulator) instead of summed
 * sequentially: the device's definition */
void mo_resample_trans_u01_rep_cr(
    size_t N, size_t M, const double *w, const double *u01, uint64_t *rep);
void mo_resample_trans_rep_index(
    size_t n, size_t m, const uint64_t *rep, uint64_t *idx);
size_t mo_resample_trans_residual(size_t N, size_t M, const double *w,
    double *resid, uint64_t *integ);
void mo_resample(int scheme, size_t n, size_t m, mo_engine *e,
    const double *w, uint64_t *rep, uint64_t *idx);

/* ---- this project's SMC (device algorithm restated on the CPU) ---- */
/* models of the SMC oracle: 1 linear-Gaussian, 2 stochastic volatility,
 * 3 the nonstationary growth model (a model outside the library, for the
 * fused samplers on a caller's own model): theta = (s, sy, sd0),
 * x_t = (0.5 x + 25 x / (1 + x^2)) + u_t + s e, y_t = x^2 / 20 + sy w, the
 * T observations followed by the T drift values u_t in y */
enum { MO_MODEL_LG = 1, MO_MODEL_SV = 2, MO_MODEL_GROWTH = 3, MO_MODEL_CV2 = 4 };
float mo_log_f(float x);
float mo_exp_f(float x);
void mo_sincos2pi_f(float u, float *s, float *c);
float mo_smc_normal(uint64_t seed, uint64_t i, uint64_t t);
double mo_smc_uniform(uint64_t seed, uint64_t t, uint64_t j);
float mo_smc_obs(int model, const double *theta, double y);
float mo_smc_logg(int model, const double *theta, float obs, float x);
double mo_smc_run(int model, const double *theta, const double *y, size_t T,
    size_t N, uint64_t seed, double threshold, int scheme, int nthreads,
    double *ess_hist, int *resampled_hist);
double mo_smc_run_est(int model, const double *theta, const double *y,
    size_t T, size_t N, uint64_t seed, double threshold, int scheme,
    int nthreads, double *ess_hist, int *resampled_hist, int kmom,
    double *est);
/* The fused sampler's resampling step, for checks against the reference's
 * own transforms: runs mo_smc_run's algorithm up to its first resampling
 * step at or after t_first (any of the six schemes) and writes that step's
 * uniform draws u01 (systematic forms: u01[0] only), the normalised
 * weights W (N), the uniform sequence u (N; residual schemes: the R
 * uniforms of the remainder) and rep (N).  Returns the step, or 0 when no
 * step from t_first on resamples. */
size_t mo_smc_resample_dump(int model, const double *theta, const double *y,
    size_t T, size_t N, uint64_t seed, double threshold, int scheme,
    size_t t_first, double *u01, double *W, double *u, uint64_t *rep);

#ifdef __cplusplus
}
#endif

#endif
