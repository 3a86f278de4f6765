/*
 * mckl_hip.h -- C ABI of the MI355X (gfx950) implementation of MCKL's Monte
 * Carlo hot path.
 *
 * Every entry point returns an int status: 0 on success, a negative value on
 * failure (MCKL_HIP_EINVAL for invalid arguments, MCKL_HIP_EHIP - code for a
 * HIP runtime error).  mckl_hip_last_error() returns a description of the
 * last failure on the calling thread.  Pointers named d_* are device memory;
 * h_* are host memory.  `stream` is a hipStream_t (NULL = default stream).
 * Calls on distinct engine states and streams are thread-safe.  Output bits
 * depend only on (key, counter, n, parameters), never on grid shape.
 *
 * Reference paths below are relative to the zhouyan/MCKL source tree.
 */
#ifndef MCKL_HIP_H
#define MCKL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCKL_HIP_OK 0
#define MCKL_HIP_EINVAL (-22)
#define MCKL_HIP_ENOMEM (-12)
#define MCKL_HIP_EHIP (-1000)

/* engine kinds */
#define MCKL_HIP_PHILOX4X32 1   /* random/philox.hpp:87-180, 10 rounds */
#define MCKL_HIP_THREEFRY4X64 2 /* random/threefry.hpp:82-205, 20 rounds */
#define MCKL_HIP_ARS 3          /* random/aes.hpp:135-214, ARS-5 */

/* U01 intervals: random/u01_distribution.hpp:140-211 */
#define MCKL_HIP_U01_CC 0
#define MCKL_HIP_U01_CO 1
#define MCKL_HIP_U01_OC 2
#define MCKL_HIP_U01_OO 3
#define MCKL_HIP_U01_CANONICAL 4

/* resampling schemes: ResampleScheme, algorithm/resample.hpp:43-50 (same
 * numbering) */
#define MCKL_HIP_MULTINOMIAL 0
#define MCKL_HIP_STRATIFIED 1
#define MCKL_HIP_SYSTEMATIC 2
#define MCKL_HIP_RESIDUAL 3
#define MCKL_HIP_RESIDUAL_STRATIFIED 4
#define MCKL_HIP_RESIDUAL_SYSTEMATIC 5

/* matrix layouts: MatrixLayout, internal/defines.hpp:46 */
#define MCKL_HIP_ROW_MAJOR 101
#define MCKL_HIP_COL_MAJOR 102

/* models of the built-in SMC samplers */
#define MCKL_HIP_MODEL_LG 1 /* x' = a x + sx e, y = x + sy n, x0 ~ N(0,1) */
#define MCKL_HIP_MODEL_SV 2 /* x' = phi x + s e, y = beta exp(x/2) n */

/*
 * Counter-engine state.  Mirrors CounterEngine (random/counter.hpp:61-288):
 * key, counter and the number of words of the buffered block already
 * consumed (index == words per block means "empty").  The buffered block is
 * the generator applied to `ctr` and is recomputed on the device when
 * needed, so the state is plain host data that the host advances exactly as
 * the reference's bulk and scalar calls advance it.
 *   Philox4x32: key[0..1] hold the two u32 key words; ctr[0..1] used.
 *   Threefry4x64: key[0..3] u64; ctr[0..3].
 *   ARS: key[0..3] hold the four u32 key words; ctr[0..1].
 * rbits is the result width (32, or 64 for the *_64 engines).
 */
typedef struct mckl_hip_engine {
    uint64_t ctr[4];
    uint64_t key[4];
    uint32_t index;
    int32_t kind;
    int32_t rbits;
    int32_t reserved;
} mckl_hip_engine;

const char *mckl_hip_last_error(void);

/* Words of result_type per block (Generator::size()/sizeof(ResultType)). */
int mckl_hip_engine_words(const mckl_hip_engine *e);

/* CounterEngine(const key_type&) / seed(key): counter.hpp:112-114,282-287 */
int mckl_hip_engine_init(
    mckl_hip_engine *e, int kind, int rbits, const uint64_t *key);

/* seed(result_type s): key = {s, 0, ...}; counter.hpp:97-103 */
int mckl_hip_engine_seed(mckl_hip_engine *e, uint64_t s);

/* ctr(const ctr_type&): counter.hpp:123-127 */
int mckl_hip_engine_set_ctr(mckl_hip_engine *e, const uint64_t *ctr);

/* discard(skip_type): counter.hpp:178-199 (host arithmetic only) */
int mckl_hip_engine_discard(mckl_hip_engine *e, uint64_t nskip);

/* Scalar operator()(): counter.hpp:134-142.  Computes the block on the
 * device and returns the word to the host (synchronous). */
int mckl_hip_engine_next(mckl_hip_engine *e, uint64_t *h_out);

/* The buffered block G(ctr) as bytes (16 or 32), computed on the device
 * (synchronous); used for the reference's textual state output. */
int mckl_hip_engine_block(const mckl_hip_engine *e, void *h_block);

/*
 * SeedGenerator<key_type>::get() called n times (random/seed.hpp:54-341) as
 * RNGSetVector::reset() uses it (random/rng_set.hpp:81-120; the seed type of
 * a counter engine is its key_type, random/counter.hpp:291-296): starting
 * from internal seed `seed` (the reference starts at 1), with partition
 * (np, rank) (seed.hpp:171-199) and Threefry randomization when `randomize`
 * (MCKL_SEED_RANDOMIZE, the reference's default).  Writes n keys of 4 u64 in
 * mckl_hip_engine.key layout to device memory.  Asynchronous.
 */
int mckl_hip_seed_keys(int kind, uint64_t seed, uint64_t np, uint64_t rank,
    int randomize, size_t n, uint64_t *d_keys, void *stream);

/*
 * Fresh per-engine streams of an RNG set (RNGSetVector<RNG>(n) after reset,
 * random/rng_set.hpp:81-120): engine e has key d_keys[4e..4e+3] and counter
 * 0, and writes its first `per` outputs to d_r[e * per ..]: rand(rng[e], per,
 * r) (counter.hpp:298-303) for 32-bit engines, or rand(rng[e],
 * U01xxDistribution<T>, per, r) (u01_distribution.hpp:79-142).
 */
int mckl_hip_rngset_rand(int kind, size_t nengines, const uint64_t *d_keys,
    size_t per, uint32_t *d_r, void *stream);
int mckl_hip_rngset_u01_f32(int kind, int interval, size_t nengines,
    const uint64_t *d_keys, size_t per, float *d_r, void *stream);
int mckl_hip_rngset_u01_f64(int kind, int interval, size_t nengines,
    const uint64_t *d_keys, size_t per, double *d_r, void *stream);

/*
 * Raw generator kernels with the shape of the reference's native C ABI
 * (include/mckl/mckl.h:313-341, e.g. mckl_philox4x32_sse2_kernel): n blocks
 * for counters ctr+1 .. ctr+n written to d_r (n * 16 or n * 32 bytes), and
 * ctr advanced by n.  Unlike the NASM kernels the counter carries across all
 * words (the reference's callers split at the carry to the same effect:
 * random/internal/philox_sse2_4x32.hpp:71-92).  Asynchronous on `stream`.
 */
int mckl_hip_philox4x32_kernel(
    uint64_t *ctr, size_t n, void *d_r, const uint32_t *key, void *stream);
int mckl_hip_threefry4x64_kernel(
    uint64_t *ctr, size_t n, void *d_r, const uint64_t *key, void *stream);
int mckl_hip_ars_kernel(
    uint64_t *ctr, size_t n, void *d_r, const uint32_t *key, void *stream);

/* rand(rng, n, r): counter.hpp:142-165,298-303.  n words of result_type
 * (u32 for rbits 32, u64 for rbits 64) into d_r; state advanced as the
 * reference's bulk call advances it. */
int mckl_hip_rand(mckl_hip_engine *e, size_t n, void *d_r, void *stream);

/* rand(rng, UniformBitsDistribution<UIntType>, n, r) with UIntType of ubits
 * = 16, 32 or 64 bits: random/uniform_bits_distribution.hpp:60-107.  Wider
 * than the engine's result (u64 from a 32-bit engine): the result words
 * themselves, little-endian, ubits/rbits per value; narrower: the low ubits
 * of one result word per value, in chunks of BufferSize<UIntType>. */
int mckl_hip_rand_uniform_bits(
    mckl_hip_engine *e, int ubits, size_t n, void *d_r, void *stream);

/* The device addresses, on the current device, of the tables that the
 * correctly rounded log, log1p, sin and cos of device evaluation objects
 * read (mckl_hip_eval.hpp's DeviceMath).  No counterpart in the reference:
 * its evaluation objects call the host libm through math/vmf.hpp; these are
 * the device stand-ins, rounded once. */
int mckl_hip_device_math(
    const double **d_log_c, const double **d_sin_a, const double **d_cos_a);

/* rand(rng, U01xxDistribution<T>, n, r): random/u01_distribution.hpp:79-142,
 * random/internal/u01_generic.hpp:48-198, random/internal/common.hpp:69-86
 * (chunking by BufferSize<T>).  32-bit engines take one u32 per value (two
 * for Canonical<double>); 64-bit-result engines one u64 per value
 * (U01UIntType, u01_distribution.hpp:105-113). */
int mckl_hip_rand_u01_f32(
    mckl_hip_engine *e, int interval, size_t n, float *d_r, void *stream);
int mckl_hip_rand_u01_f64(
    mckl_hip_engine *e, int interval, size_t n, double *d_r, void *stream);

/* rand(rng, NormalDistribution<T>(mean, sd), n, r):
 * random/normal_distribution.hpp:50-101 (chunked Box-Muller, odd tail), on
 * U01OC values of one result word each (u32 or u64).  log, sin and cos are
 * the correctly rounded values. */
int mckl_hip_rand_normal_f32(mckl_hip_engine *e, size_t n, float mean,
    float sd, float *d_r, void *stream);
int mckl_hip_rand_normal_f64(mckl_hip_engine *e, size_t n, double mean,
    double sd, double *d_r, void *stream);

/* rand(rng, UniformRealDistribution<T>(a, b), n, r), either result width:
 * random/uniform_real_distribution.hpp:48-58 (U01CO, then r*(b-a)+a rounded
 * twice, the reference's muladd without FMA, math/vmf.hpp:697-722); a < b
 * (:43-46) or MCKL_HIP_EINVAL. */
int mckl_hip_rand_uniform_real_f32(mckl_hip_engine *e, size_t n, float a,
    float b, float *d_r, void *stream);
int mckl_hip_rand_uniform_real_f64(mckl_hip_engine *e, size_t n, double a,
    double b, double *d_r, void *stream);

/* The same fills for `nstreams` independent engines in one launch: engine s
 * writes n values to d_r + s*n and is advanced as if it had made the
 * single-engine call.  (No reference counterpart: the reference loops over
 * engines; one launch keeps the GPU full.) */
int mckl_hip_rand_u01_f64_streams(mckl_hip_engine *e, size_t nstreams,
    int interval, size_t n, double *d_r, void *stream);
int mckl_hip_rand_normal_f32_streams(mckl_hip_engine *e, size_t nstreams,
    size_t n, float mean, float sd, float *d_r, void *stream);
int mckl_hip_rand_normal_f64_streams(mckl_hip_engine *e, size_t nstreams,
    size_t n, double mean, double sd, double *d_r, void *stream);

/*
 * Weight (core/weight.hpp:44-228) on device-resident normalised FP64
 * weights d_w.  d_stats (device, 2 doubles) receives {ESS, lmax+log(accw)};
 * the second value is the log normaliser increment that the reference
 * computes in normalize() and discards (weight.hpp:198-227).
 */
int mckl_hip_weight_set_equal(size_t n, double *d_w, double *d_stats,
    void *stream); /* weight.hpp:76-80 */
int mckl_hip_weight_set_log(size_t n, const double *d_v, double *d_w,
    double *d_stats, void *stream); /* weight.hpp:124-129 */
int mckl_hip_weight_add_log(size_t n, const double *d_v, double *d_w,
    double *d_stats, void *stream); /* weight.hpp:143-148 */

/*
 * Resampling (algorithm/resample.hpp).  Replication counts and parent
 * indices are u32.  d_u01: M sorted uniforms.
 */
/* u01_rand_sorted / _stratified / _systematic: resample.hpp:216-268 */
int mckl_hip_u01_rand_sequence(
    int scheme, mckl_hip_engine *e, size_t n, double *d_r, void *stream);
/* resample_trans_u01_rep: resample.hpp:374-407 */
int mckl_hip_resample_trans_u01_rep(size_t N, size_t M, const double *d_w,
    const double *d_u01, uint32_t *d_rep, void *stream);
/* resample_trans_rep_index: resample.hpp:416-466 */
int mckl_hip_resample_trans_rep_index(size_t N, size_t M,
    const uint32_t *d_rep, uint32_t *d_idx, void *stream);
/* ResampleAlgorithm::operator() + rep->index: resample.hpp:515-576,
 * ResampleEval::operator() resample.hpp:494-507 */
int mckl_hip_resample(int scheme, size_t N, size_t M, mckl_hip_engine *e,
    const double *d_w, uint32_t *d_rep, uint32_t *d_idx, void *stream);
/* No counterpart in the reference.  mckl_hip_resample for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops): the same results and engine advance, with the workspace taken from
 * a device array of the library (about 36 MiB, held by each device the
 * library runs on until the process ends) instead of the stream-ordered
 * allocator.  N, M <= 2^20 (1048576), else MCKL_HIP_EINVAL.  Every call on
 * one device shares that array: calls that overlap on one device (other
 * streams, other threads) race and must be serialised by the caller.  The
 * residual schemes read their residual count back to the host (a stream
 * synchronisation), as mckl_hip_resample does. */
int mckl_hip_resample_noalloc(int scheme, size_t N, size_t M,
    mckl_hip_engine *e, const double *d_w, uint32_t *d_rep, uint32_t *d_idx,
    void *stream);
/* SMCEstimator::estimate's weighted sums (algorithm/smc.hpp:89-118, a dgemv
 * in the reference): d_r[j] = sum_i d_w[i] u(i, j), j < dim, with u(i, j) =
 * d_u[i*dim + j] (MCKL_HIP_ROW_MAJOR) or d_u[j*n + i] (MCKL_HIP_COL_MAJOR);
 * the f32 form converts u to double first, as SMCEstimator<T, float> does
 * (smc.hpp:128-135).  Sums are double-double, so d_r[j] is the exact sum
 * rounded once, to ~2^-104 relative to sum |w u|.  d_r is device memory. */
int mckl_hip_estimate_f64(size_t n, size_t dim, int layout, const double *d_u,
    const double *d_w, double *d_r, void *stream);
int mckl_hip_estimate_f32(size_t n, size_t dim, int layout, const float *d_u,
    const double *d_w, double *d_r, void *stream);

/* StateMatrix<RowMajor,dim,T>::select: core/state_matrix.hpp:193-265
 * (out of place: d_dst[j] = d_src[d_idx[j]], rows of `dim` values; the
 * indices must address rows of d_src) */
int mckl_hip_select_f32(size_t M, size_t dim, const uint32_t *d_idx,
    const float *d_src, float *d_dst, void *stream);
int mckl_hip_select_f64(size_t M, size_t dim, const uint32_t *d_idx,
    const double *d_src, double *d_dst, void *stream);
/* StateMatrix<ColMajor,dim,T>::select: core/state_matrix.hpp:240-265 (out
 * of place: column c of d_dst, M rows, is column c of d_src, N rows,
 * gathered by d_idx; the indices must lie below N) */
int mckl_hip_select_cols_f32(size_t M, size_t N, size_t dim,
    const uint32_t *d_idx, const float *d_src, float *d_dst, void *stream);
int mckl_hip_select_cols_f64(size_t M, size_t N, size_t dim,
    const uint32_t *d_idx, const double *d_src, double *d_dst, void *stream);

/*
 * ResampleIndex<uint32_t> (algorithm/resample.hpp:597-832): the history of
 * resampling indices, one device vector per iteration, and its trace-back.
 * Bookkeeping follows the reference's member by member: reset() zeroes the
 * iteration count and keeps the vectors for reuse, clear() frees them and
 * keeps the count, push_back() stores iteration num_iter() into vector
 * num_iter()-1 when that vector exists and appends a vector otherwise.
 * Queries the reference guards with runtime_assert return MCKL_HIP_EINVAL,
 * and so do queries that would read a vector clear() has freed.  An entry
 * that does not index the vector it is read against traces to 0xFFFFFFFF.
 * Calls on one handle are not thread-safe.
 */
typedef struct mckl_hip_resample_index mckl_hip_resample_index;
int mckl_hip_resample_index_create(mckl_hip_resample_index **out);
int mckl_hip_resample_index_destroy(mckl_hip_resample_index *h);
/* num_iter(): resample.hpp:606 */
int mckl_hip_resample_index_num_iter(
    const mckl_hip_resample_index *h, size_t *out);
/* size() (iter = SIZE_MAX) and size(iter): resample.hpp:609-625 */
int mckl_hip_resample_index_size(
    const mckl_hip_resample_index *h, size_t iter, size_t *out);
/* reset(), clear(): resample.hpp:628-631 */
int mckl_hip_resample_index_reset(mckl_hip_resample_index *h);
int mckl_hip_resample_index_clear(mckl_hip_resample_index *h);
/* push_back(N) (d_idx NULL: identity) and push_back(N, first):
 * resample.hpp:634-656; d_idx is copied on `stream` */
int mckl_hip_resample_index_push_back(mckl_hip_resample_index *h, size_t N,
    const uint32_t *d_idx, void *stream);
/* index(id, iter_back, iter) for n ids at once: resample.hpp:658-681.
 * As in the reference the trace starts from the last vector held (not from
 * vector iter_back) and then walks vectors iter_back-1 down to iter. */
int mckl_hip_resample_index_index(const mckl_hip_resample_index *h,
    size_t n, const uint32_t *d_id, size_t iter_back, size_t iter,
    uint32_t *d_out, void *stream);
/* index_matrix_nrow / ncol: resample.hpp:683-715 */
int mckl_hip_resample_index_matrix_shape(const mckl_hip_resample_index *h,
    size_t iter_back, size_t iter, size_t *nrow, size_t *ncol);
/* read_index_matrix(layout, first, iter_back, iter): resample.hpp:741-810;
 * d_out holds nrow * ncol entries, element (i, c) = trace of particle i of
 * iteration iter_back back to iteration iter + c */
int mckl_hip_resample_index_read_matrix(const mckl_hip_resample_index *h,
    int layout, size_t iter_back, size_t iter, uint32_t *d_out, void *stream);

/*
 * Plain SMC sampler (SMCSampler::iterate, algorithm/smc.hpp:352-406) for the
 * built-in models, FP32 states and log-weights, FP64 reductions and scans.
 * The handle owns its device buffers; run() is asynchronous and contains no
 * host synchronisation (the ESS test is taken on the device), so it can be
 * captured into a HIP graph.  Schemes (SMCSampler::resample_method,
 * smc.hpp:260-282): all six, the three plain schemes and the three residual
 * ones (integral parts kept, R residual uniforms of the base scheme).
 */
typedef struct mckl_hip_smc mckl_hip_smc;
int mckl_hip_smc_create(int model, const double *theta, size_t T,
    const double *h_y, size_t N, uint64_t seed, double threshold, int scheme,
    mckl_hip_smc **out);
int mckl_hip_smc_run(mckl_hip_smc *s, void *stream);
/* loglik, and optionally per-step ESS and resampling flags (length T) */
int mckl_hip_smc_result(mckl_hip_smc *s, double *h_loglik, double *h_ess,
    int32_t *h_resampled);
int mckl_hip_smc_destroy(mckl_hip_smc *s);
/* Selection estimator (SMCEstimator at step 0, algorithm/smc.hpp:89-118,
 * :389-394) of the moments x^1..x^k, 0 <= k <= 4 (0 removes it): after the
 * move and weighting of each step, est[t][j-1] = sum_i W_i x_i^j with the
 * normalised weights, computed in double-double from the FP32 states and
 * weights of the run.  Call before run(); estimates() returns T x k values
 * (row-major, one row per step) and synchronises. */
int mckl_hip_smc_estimate_moments(mckl_hip_smc *s, int k);
int mckl_hip_smc_estimates(mckl_hip_smc *s, double *h_est);

/*
 * Batched SMC: nfilters independent filters of uneven sizes in one
 * persistent launch (no reference counterpart; the reference runs one
 * SMCSampler at a time).  Filter f uses theta[3f..3f+2], y[T f .. T f+T-1],
 * seed[f] and is bit-identical to the plain sampler with the same inputs.
 * All six schemes; at most 131072 particles per filter.
 */
typedef struct mckl_hip_smc_batch mckl_hip_smc_batch;
/* Launch configuration of a batched run.  No field changes a bit of the
 * results.  split_tiles > 0 runs the filters of at least split_tiles
 * 2048-particle tiles in a persistent launch of `block`-thread workgroups
 * and the others concurrently in one of `small_block`-thread workgroups;
 * 0 runs every filter in one launch of `block`. */
typedef struct mckl_hip_tune {
    int32_t block;        /* threads per workgroup: 256, 512 or 1024 */
    int32_t waves_per_eu; /* occupancy cap: 0 = as many as fit */
    int32_t split_tiles;  /* 0 = one launch */
    int32_t small_block;  /* threads per workgroup of the small filters */
} mckl_hip_tune;
int mckl_hip_smc_batch_create(int model, size_t nfilters,
    const uint32_t *h_N, const double *h_theta, size_t T, const double *h_y,
    const uint64_t *h_seed, double threshold, int scheme,
    mckl_hip_smc_batch **out);
int mckl_hip_smc_batch_set_tune(mckl_hip_smc_batch *b, const mckl_hip_tune *t);
int mckl_hip_smc_batch_run(mckl_hip_smc_batch *b, int use_graph, void *stream);
/* Autotuner (no reference counterpart): times a fixed list of launch
 * configurations (workgroup width, occupancy cap, large/small split) with
 * graph runs on `stream`, checks that each reproduces the first one's
 * log-likelihoods bit for bit (MCKL_HIP_EHIP otherwise), and keeps the
 * fastest.  With cache_path, a configuration already recorded there for the
 * same workload signature (device, model, scheme, T, threshold, filter
 * sizes) is taken without runs, and a new choice is appended (one JSON
 * object per line).  best and best_ms (ms per run) may be NULL. */
int mckl_hip_smc_batch_autotune(mckl_hip_smc_batch *b, const char *cache_path,
    mckl_hip_tune *best, double *best_ms, void *stream);
int mckl_hip_smc_batch_result(mckl_hip_smc_batch *b, double *h_loglik);
/* New parameters theta[3f..3f+2] and seeds for every filter of an existing
 * batch (same model, sizes, observations, threshold and scheme); the next
 * run uses them, with or without the captured graph.  Waits for runs in
 * flight.  Used by the PMCMC drivers, which re-run the same filters at
 * proposed parameters (algorithm/pmcmc.hpp:138-142). */
int mckl_hip_smc_batch_set_params(mckl_hip_smc_batch *b,
    const double *h_theta, const uint64_t *h_seed);
int mckl_hip_smc_batch_destroy(mckl_hip_smc_batch *b);

/*
 * Gradient of the SMC log-likelihood estimate with respect to theta by a
 * hand-written adjoint (no reference counterpart: the reference exposes
 * neither the log-likelihood, core/weight.hpp:198-227, nor derivatives).
 * nbatch parameter points, N <= 131072 particles each, FP64 states; the
 * ancestors and resampling decisions are those of the forward pass at theta
 * and stay fixed in the reverse sweep; forward states are checkpointed every
 * K steps and recomputed segment by segment.  y holds T observations shared
 * by all points (y_per_point = 0) or T per point.  theta, grad: 3 per point
 * (LG: a, sx, sy; SV: phi, sigma, beta).  Schemes: all six (the three base
 * schemes and their residual forms, resample.hpp:581-593).
 * With tape = 1 run() also records the ancestors, normalised log-weights and
 * decisions (nbatch * (T+1) * N * 12 bytes), and replay() evaluates the
 * differentiated function at another theta with them held fixed: the
 * function whose central differences the gradient must match.
 */
typedef struct mckl_hip_smc_grad mckl_hip_smc_grad;
int mckl_hip_smc_grad_create(int model, size_t nbatch, size_t N, size_t T,
    const double *h_y, int y_per_point, const uint64_t *h_seed,
    double threshold, int scheme, size_t K, int tape,
    mckl_hip_smc_grad **out);
/* forward + adjoint: h_loglik[nbatch], h_grad[3 nbatch] (either may be NULL);
 * synchronises the stream */
int mckl_hip_smc_grad_run(mckl_hip_smc_grad *g, const double *h_theta,
    double *h_loglik, double *h_grad, void *stream);
int mckl_hip_smc_grad_replay(mckl_hip_smc_grad *g, const double *h_theta,
    double *h_value, void *stream);
int mckl_hip_smc_grad_destroy(mckl_hip_smc_grad *g);
/* one-shot form: create, run, destroy (no tape) */
int mckl_hip_smc_loglik_grad(int model, size_t nbatch, const double *h_theta,
    size_t N, size_t T, const double *h_y, int y_per_point,
    const uint64_t *h_seed, double threshold, int scheme, size_t K,
    double *h_loglik, double *h_grad, void *stream);

#ifdef __cplusplus
}
#endif

#endif
