// resample_index.hip -- ResampleIndex (algorithm/resample.hpp:597-832) with
// the history resident in device memory.
//
// The trace-back of read_index_matrix is a chain of gathers: column R-1 is
// vector iter_back itself and column c is vector iter+c gathered at column
// c+1.  A thread follows two particles of iteration iter_back through up to
// kChain vectors per launch (their pointers travel as kernel arguments) and
// keeps the running indices in registers, so two independent gathers are in
// flight per step; a longer chain takes further launches that restart from
// the last column written.  Column-major output is stored directly (lanes
// write consecutive particles).  Row-major output is staged per block in
// LDS and written out row segment by row segment, so a wave still writes
// contiguous runs instead of one word per row.
#include "common.hpp"

#include <algorithm>
#include <vector>

struct mckl_hip_resample_index {
    std::vector<uint32_t *> vec; // index_[k]
    std::vector<size_t> len;     // index_[k].size()
    std::vector<size_t> cap;     // allocated entries of vec[k]
    size_t num_iter = 0;         // num_iter_
};

namespace mckl_hip {
namespace {

constexpr uint32_t kBad = 0xFFFFFFFFu;
constexpr int kChain = 64;
constexpr int kRB = 256; // threads per block

struct Chain {
    const uint32_t *p[kChain];
    uint32_t n[kChain];
};

__global__ __launch_bounds__(kRB) void k_iota(uint64_t n, uint32_t *out)
{
    const uint64_t stride = (uint64_t)gridDim.x * kRB;
    for (uint64_t i = (uint64_t)blockIdx.x * kRB + threadIdx.x; i < n;
         i += stride)
        out[i] = (uint32_t)i;
}

// out[j] = src[j] pushed through the cnt vectors of c, in order
__global__ __launch_bounds__(kRB) void k_trace_ids(uint64_t n,
    const uint32_t *src, int cnt, Chain c, uint32_t *out)
{
    const uint64_t j = (uint64_t)blockIdx.x * kRB + threadIdx.x;
    if (j >= n)
        return;
    uint32_t a = src[j];
    for (int k = 0; k < cnt; ++k)
        a = a < c.n[k] ? c.p[k][a] : kBad;
    out[j] = a;
}

// Columns col0, col0-1, ..., col0-cnt+1 of the N x R index matrix.  Each
// thread follows kP particles (base + p * kRB + tid) so that kP independent
// gathers are in flight per step of the chain.  The running index starts at
// the particle number (first launch) or at column col0+1 written by the
// previous launch.  Row-major output goes through an LDS tile of kSub
// columns per pass: rows of the tile are contiguous kSub-word runs of the
// output rows.
template <bool kRow, int kP, int kSub>
__global__ __launch_bounds__(kRB) void k_trace_matrix(uint32_t N, uint64_t R,
    uint64_t col0, int cnt, int first, Chain c, uint32_t *__restrict__ out)
{
    __shared__ uint32_t tile[kRow ? kP * kRB * (kSub + 1) : 1];
    const uint32_t base = blockIdx.x * (uint32_t)(kP * kRB);
    uint32_t a[kP];
    bool ok[kP];
#pragma unroll
    for (int p = 0; p < kP; ++p) {
        const uint32_t i = base + p * kRB + threadIdx.x;
        ok[p] = i < N;
        a[p] = first ? i
                     : !ok[p] ? 0u
                              : kRow ? out[(uint64_t)i * R + col0 + 1]
                                     : out[(col0 + 1) * N + i];
    }
    if (!kRow) {
        for (int k = 0; k < cnt; ++k) {
            const uint32_t *v = c.p[k];
            const uint32_t n = c.n[k];
            uint32_t b[kP];
#pragma unroll
            for (int p = 0; p < kP; ++p)
                b[p] = ok[p] && a[p] < n ? v[a[p]] : kBad;
#pragma unroll
            for (int p = 0; p < kP; ++p) {
                a[p] = b[p];
                if (ok[p])
                    out[(col0 - k) * N + base + p * kRB + threadIdx.x] = a[p];
            }
        }
        return;
    }
    const uint32_t rows = min((uint32_t)(kP * kRB), N - base);
    for (int s = 0; s < cnt; s += kSub) {
        const int m = min(kSub, cnt - s);
        for (int k = 0; k < m; ++k) {
            const uint32_t *v = c.p[s + k];
            const uint32_t n = c.n[s + k];
            uint32_t b[kP];
#pragma unroll
            for (int p = 0; p < kP; ++p)
                b[p] = ok[p] && a[p] < n ? v[a[p]] : kBad;
#pragma unroll
            for (int p = 0; p < kP; ++p) {
                a[p] = b[p];
                tile[(p * kRB + threadIdx.x) * (kSub + 1) + k] = a[p];
            }
        }
        __syncthreads();
        // entries lo .. lo+m-1 = col0-s-m+1 .. col0-s of each row
        const uint64_t lo = col0 - s + 1 - m;
        const uint32_t total = rows * (uint32_t)m;
        for (uint32_t e = threadIdx.x; e < total; e += kRB) {
            const uint32_t r = e / (uint32_t)m;
            const uint32_t q = e - r * (uint32_t)m;
            out[(uint64_t)(base + r) * R + lo + q] =
                tile[r * (kSub + 1) + (m - 1 - q)];
        }
        __syncthreads();
    }
}

// two chains per thread; row-major tiles of 16 columns (34 KB of LDS)
constexpr int kPCol = 2, kPRow = 2, kSubRow = 16;

int fail_range(const char *what)
{
    return set_error(MCKL_HIP_EINVAL,
        "**ResampleIndex::%s** iteration numbers out of range", what);
}

// vectors iter_back-1 .. iter must exist (clear() frees them)
int check_held(const mckl_hip_resample_index *h, size_t iter_back)
{
    if (iter_back >= h->vec.size())
        return set_error(MCKL_HIP_EINVAL,
            "ResampleIndex: iteration %zu was released by clear()", iter_back);
    return MCKL_HIP_OK;
}

void fill_chain(const mckl_hip_resample_index *h, Chain &c, int k, size_t v)
{
    c.p[k] = h->vec[v];
    c.n[k] = (uint32_t)h->len[v];
}

} // namespace
} // namespace mckl_hip

using namespace mckl_hip;

extern "C" {

int mckl_hip_resample_index_create(mckl_hip_resample_index **out)
{
    if (!out)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    *out = new mckl_hip_resample_index();
    return MCKL_HIP_OK;
}

int mckl_hip_resample_index_clear(mckl_hip_resample_index *h)
{
    if (!h)
        return set_error(MCKL_HIP_EINVAL, "NULL handle");
    int rc = MCKL_HIP_OK;
    for (uint32_t *p : h->vec)
        if (p && rc == MCKL_HIP_OK)
            rc = hip_status(hipFree(p), "hipFree");
    h->vec.clear();
    h->len.clear();
    h->cap.clear();
    return rc;
}

int mckl_hip_resample_index_destroy(mckl_hip_resample_index *h)
{
    if (!h)
        return MCKL_HIP_OK;
    const int rc = mckl_hip_resample_index_clear(h);
    delete h;
    return rc;
}

int mckl_hip_resample_index_num_iter(
    const mckl_hip_resample_index *h, size_t *out)
{
    if (!h || !out)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    *out = h->num_iter;
    return MCKL_HIP_OK;
}

int mckl_hip_resample_index_size(
    const mckl_hip_resample_index *h, size_t iter, size_t *out)
{
    if (!h || !out)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (iter == SIZE_MAX)
        *out = h->len.empty() ? 0 : h->len.back();
    else
        *out = iter < h->len.size() ? h->len[iter] : 0;
    return MCKL_HIP_OK;
}

int mckl_hip_resample_index_reset(mckl_hip_resample_index *h)
{
    if (!h)
        return set_error(MCKL_HIP_EINVAL, "NULL handle");
    h->num_iter = 0;
    return MCKL_HIP_OK;
}

int mckl_hip_resample_index_push_back(mckl_hip_resample_index *h, size_t N,
    const uint32_t *d_idx, void *stream)
{
    if (!h)
        return set_error(MCKL_HIP_EINVAL, "NULL handle");
    if (N > (size_t)kBad)
        return set_error(MCKL_HIP_EINVAL,
            "ResampleIndex: sample size %zu exceeds 2^32 - 1", N);
    size_t k;
    if (h->vec.size() < h->num_iter + 1) {
        k = h->vec.size();
        h->vec.push_back(nullptr);
        h->len.push_back(0);
        h->cap.push_back(0);
    } else {
        k = h->num_iter;
    }
    if (h->cap[k] < N) {
        if (h->vec[k])
            MCKL_HIP_TRY(hipFree(h->vec[k]));
        h->vec[k] = nullptr;
        h->cap[k] = 0;
        void *p = nullptr;
        MCKL_HIP_TRY(hipMalloc(&p, N * sizeof(uint32_t)));
        h->vec[k] = static_cast<uint32_t *>(p);
        h->cap[k] = N;
    }
    h->len[k] = N;
    ++h->num_iter;
    if (N == 0)
        return MCKL_HIP_OK;
    hipStream_t st = as_stream(stream);
    if (d_idx) {
        MCKL_HIP_TRY(hipMemcpyAsync(h->vec[k], d_idx, N * sizeof(uint32_t),
            hipMemcpyDeviceToDevice, st));
    } else {
        const uint32_t g = (uint32_t)std::min<uint64_t>(ceil_div(N, kRB), 8192);
        k_iota<<<g, kRB, 0, st>>>(N, h->vec[k]);
        MCKL_HIP_CHECK_LAUNCH("k_iota");
    }
    return MCKL_HIP_OK;
}

int mckl_hip_resample_index_index(const mckl_hip_resample_index *h,
    size_t n, const uint32_t *d_id, size_t iter_back, size_t iter,
    uint32_t *d_out, void *stream)
{
    if (!h)
        return set_error(MCKL_HIP_EINVAL, "NULL handle");
    if (!(iter <= iter_back && iter_back < h->num_iter))
        return fail_range("index");
    if (h->vec.empty())
        return check_held(h, iter_back);
    if (iter_back > iter) {
        const int rc = check_held(h, iter_back - 1);
        if (rc)
            return rc;
    }
    if (n == 0)
        return MCKL_HIP_OK;
    if (!d_id || !d_out)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    hipStream_t st = as_stream(stream);
    // the chain: the last vector held, then iter_back-1 down to iter
    std::vector<size_t> vs;
    vs.push_back(h->vec.size() - 1);
    for (size_t v = iter_back; v > iter; --v)
        vs.push_back(v - 1);
    const uint32_t g = (uint32_t)ceil_div(n, kRB);
    for (size_t s = 0; s < vs.size(); s += kChain) {
        Chain c{};
        const int cnt = (int)std::min<size_t>(kChain, vs.size() - s);
        for (int k = 0; k < cnt; ++k)
            fill_chain(h, c, k, vs[s + k]);
        k_trace_ids<<<g, kRB, 0, st>>>(n, s == 0 ? d_id : d_out, cnt, c, d_out);
        MCKL_HIP_CHECK_LAUNCH("k_trace_ids");
    }
    return MCKL_HIP_OK;
}

int mckl_hip_resample_index_matrix_shape(const mckl_hip_resample_index *h,
    size_t iter_back, size_t iter, size_t *nrow, size_t *ncol)
{
    if (!h || !nrow || !ncol)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    if (!(iter <= iter_back && iter_back < h->num_iter))
        return fail_range("index_matrix");
    const int rc = check_held(h, iter_back);
    if (rc)
        return rc;
    *nrow = h->len[iter_back];
    *ncol = iter_back - iter + 1;
    return MCKL_HIP_OK;
}

int mckl_hip_resample_index_read_matrix(const mckl_hip_resample_index *h,
    int layout, size_t iter_back, size_t iter, uint32_t *d_out, void *stream)
{
    if (!h)
        return set_error(MCKL_HIP_EINVAL, "NULL handle");
    if (layout != MCKL_HIP_ROW_MAJOR && layout != MCKL_HIP_COL_MAJOR)
        return set_error(MCKL_HIP_EINVAL, "unknown matrix layout %d", layout);
    size_t N = 0, R = 0;
    const int rc = mckl_hip_resample_index_matrix_shape(h, iter_back, iter, &N, &R);
    if (rc)
        return rc;
    if (N == 0)
        return MCKL_HIP_OK;
    if (!d_out)
        return set_error(MCKL_HIP_EINVAL, "NULL argument");
    hipStream_t st = as_stream(stream);
    const uint32_t gc = (uint32_t)ceil_div(N, kPCol * kRB);
    const uint32_t gr = (uint32_t)ceil_div(N, kPRow * kRB);
    // column R-1-j comes from vector iter_back-j
    for (size_t j = 0; j < R; j += kChain) {
        Chain c{};
        const int cnt = (int)std::min<size_t>(kChain, R - j);
        for (int k = 0; k < cnt; ++k)
            fill_chain(h, c, k, iter_back - j - k);
        const uint64_t col0 = R - 1 - j;
        if (layout == MCKL_HIP_ROW_MAJOR)
            k_trace_matrix<true, kPRow, kSubRow><<<gr, kRB, 0, st>>>(
                (uint32_t)N, R, col0, cnt, j == 0, c, d_out);
        else
            k_trace_matrix<false, kPCol, 1><<<gc, kRB, 0, st>>>(
                (uint32_t)N, R, col0, cnt, j == 0, c, d_out);
        MCKL_HIP_CHECK_LAUNCH("k_trace_matrix");
    }
    return MCKL_HIP_OK;
}

} // extern "C"
