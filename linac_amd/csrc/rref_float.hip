// Gauss-Jordan RREF in float64 and complex128 (field_characteristic == 0).
//
// Replaces the REAL / complex branches of linac/row_reduce.cu (SetRowScales
// and RescaleRows :201-270, ThreadsReduceToMaxIndex :287-342 with |x|,
// CompareHeadToTolerance :404-418 with tolerance 1e-9, RescaleRow :117-136,
// RowReduce :138-196) driven by linac/pycuda_row_reduce.py:87-131:
//   1. every row is divided once by its largest |x| (rows that are all zero
//      are left alone instead of turning into NaN),
//   2. per column: the unused row with the largest |x| (first such row on
//      ties) is the pivot if |x| > tol; every other row r gets
//      A[r, c] -= (A[r, j] / h) A[piv, c] for c > j and A[r, j] = 0,
//   3. pivot rows are divided by their head once at the end and gathered in
//      column order (no row swaps during the sweep, as in the GF(p) path).
// The arithmetic is float64 throughout (complex: pairs of float64).
#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

#include "float_blocked.hpp"
#include "float_ops.hpp"
#include "float_refine.hpp"
#include "runtime.hpp"

namespace linac {

template <class T>
int float_blocked_run(T* A, int64_t ld, int64_t N, int64_t M, double tol, hipStream_t s, int64_t* rank,
                      uint8_t* used, int64_t* piv_rows, int64_t* piv_cols, DeviceContext* ctx, Profiler* prof,
                      FloatFactors* keep, bool f32_update);

// Refinement passes run with LINAC_REFINE; the mixed method refines until a
// pass corrects less than MIXED_STOP (relative to max |X|), at most
// MIXED_PASSES times.
constexpr int REFINE_PASSES = 2;
constexpr int MIXED_PASSES = 10;
constexpr double MIXED_STOP = 0x1p-51;

struct FloatState {
    int64_t prow;
    int64_t rank;
    CplxF factor_unused;
    int64_t pad[4];
};

// Each row divided by its max |x| (one workgroup per row).
template <class T>
__global__ __launch_bounds__(256) void k_row_scale(T* __restrict__ A, int64_t ld, int64_t N, int64_t M) {
    __shared__ double s_max[256];
    const int64_t r = blockIdx.x;
    T* row = A + r * ld;
    double m = 0.0;
    for (int64_t c = threadIdx.x; c < M; c += 256) m = fmax(m, FloatOps<T>::abs(row[c]));
    s_max[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + w]);
        __syncthreads();
    }
    const double sc = s_max[0];
    if (sc == 0.0) return;
    for (int64_t c = threadIdx.x; c < M; c += 256) row[c] = FloatOps<T>::scale(row[c], sc);
}

__global__ void k_float_init(FloatState* st, uint8_t* used, int64_t N) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x == 0) {
        st->prow = -1;
        st->rank = 0;
    }
    for (int64_t r = x; r < N; r += (int64_t)gridDim.x * blockDim.x) used[r] = 0;
}

// Pivot of column j: argmax |A[r, j]| over unused rows (smallest row on ties),
// accepted iff the maximum exceeds tol.
template <class T>
__global__ __launch_bounds__(1024) void k_float_select(const T* __restrict__ A, int64_t ld, int64_t N, int64_t j,
                                                       double tol, FloatState* st, uint8_t* used,
                                                       int64_t* piv_rows, int64_t* piv_cols) {
    __shared__ double s_v[1024];
    __shared__ int64_t s_r[1024];
    double best = -1.0;
    int64_t br = -1;
    for (int64_t r = threadIdx.x; r < N; r += 1024) {
        if (used[r]) continue;
        const double v = FloatOps<T>::abs(A[r * ld + j]);
        if (v > best) {
            best = v;
            br = r;
        }
    }
    s_v[threadIdx.x] = best;
    s_r[threadIdx.x] = br;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double v2 = s_v[threadIdx.x + w];
            const int64_t r2 = s_r[threadIdx.x + w];
            const double v1 = s_v[threadIdx.x];
            const int64_t r1 = s_r[threadIdx.x];
            if (v2 > v1 || (v2 == v1 && r2 >= 0 && (r1 < 0 || r2 < r1))) {
                s_v[threadIdx.x] = v2;
                s_r[threadIdx.x] = r2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (s_r[0] >= 0 && s_v[0] > tol) {
            const int64_t r = s_r[0];
            st->prow = r;
            used[r] = 1;
            const int64_t k = st->rank;
            piv_rows[k] = r;
            piv_cols[k] = j;
            st->rank = k + 1;
        } else {
            st->prow = -1;
        }
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_float_elim(T* __restrict__ A, int64_t ld, int64_t N, int64_t M, int64_t j,
                                                    const FloatState* __restrict__ st, int log2G) {
    const int64_t prow = st->prow;
    if (prow < 0) return;
    const int G = 1 << log2G;
    const int64_t r = ((int64_t)blockIdx.x << (8 - log2G)) + (threadIdx.x >> log2G);
    const int t = threadIdx.x & (G - 1);
    if (r >= N || r == prow) return;
    T* row = A + r * ld;
    const T a = row[j];
    if (FloatOps<T>::is_zero(a)) return;
    const T* pr = A + prow * ld;
    const T f = FloatOps<T>::div(a, pr[j]);
    for (int64_t c = j + 1 + t; c < M; c += G) row[c] = FloatOps<T>::fms(row[c], f, pr[c]);
    if (t == 0) row[j] = FloatOps<T>::zero();
}

// Row k of the result = k-th pivot row / its head; rows >= rank are zero.
template <class T>
__global__ __launch_bounds__(256) void k_float_finalize(const T* A, int64_t ld, T* out, int64_t N, int64_t M,
                                                        const FloatState* __restrict__ st,
                                                        const int64_t* __restrict__ piv_rows,
                                                        const int64_t* __restrict__ piv_cols) {
    const int64_t k = blockIdx.x;
    T* o = out + k * ld;
    if (k >= st->rank) {
        for (int64_t c = threadIdx.x; c < M; c += 256) o[c] = FloatOps<T>::zero();
        return;
    }
    const T* s = A + piv_rows[k] * ld;
    const T h = s[piv_cols[k]];
    for (int64_t c = threadIdx.x; c < M; c += 256) o[c] = FloatOps<T>::div(s[c], h);
}

template <class T>
static int rref_float_entry(T* a, int64_t rows, int64_t cols, int64_t ld, double tol, uint32_t flags,
                            void* stream_arg, int64_t* pivots_out, int64_t* rank_out) {
    if (rows < 0 || cols < 0) return fail(LINAC_ERR_INVALID, "negative shape");
    if (ld < cols) return fail(LINAC_ERR_INVALID, "ld (%lld) < cols (%lld)", (long long)ld, (long long)cols);
    if (!(tol >= 0.0)) return fail(LINAC_ERR_INVALID, "tolerance must be >= 0");
    if (rank_out) *rank_out = 0;
    if (rows == 0 || cols == 0) return LINAC_OK;
    if (!a) return fail(LINAC_ERR_INVALID, "matrix pointer is NULL");
    const bool on_device = (flags & LINAC_DEVICE_PTR) != 0;
    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipStream_t s = stream_arg ? (hipStream_t)stream_arg : ctx->stream;
    Profiler* prof = (flags & LINAC_PROFILE) ? &thread_profiler() : nullptr;

    T* A = a;
    int64_t lda = ld;
    void* p = nullptr;
    if (!on_device) {
        lda = round_up(cols, 32);
        LINAC_TRY(ctx->get(SLOT_MATRIX, (size_t)rows * lda * sizeof(T), &p));
        A = (T*)p;
        LINAC_HIP(hipMemcpy2DAsync(A, lda * sizeof(T), a, ld * sizeof(T), cols * sizeof(T), rows,
                                   hipMemcpyHostToDevice, s));
    }
    const int64_t cap = std::max<int64_t>(1, std::min(rows, cols));
    LINAC_TRY(ctx->get(SLOT_STATE, 256 + 2 * round_up(cap * 8, 256) + round_up(rows, 256), &p));
    FloatState* st = (FloatState*)p;
    int64_t* piv_rows = (int64_t*)((char*)p + 256);
    int64_t* piv_cols = (int64_t*)((char*)p + 256 + round_up(cap * 8, 256));
    uint8_t* used = (uint8_t*)((char*)p + 256 + 2 * round_up(cap * 8, 256));
    const uint32_t method = flags & LINAC_METHOD_MASK;
    if (method != LINAC_METHOD_PLAIN && method != LINAC_METHOD_AUTO && method != LINAC_METHOD_BLOCKED &&
        method != LINAC_METHOD_MIXED)
        return fail(LINAC_ERR_INVALID, "unknown method flag 0x%x", method);
    const bool mixed = method == LINAC_METHOD_MIXED;
    if (mixed && !std::is_same<T, double>::value)
        return fail(LINAC_ERR_INVALID, "the mixed method is for float64 matrices");
    const bool refine = (flags & LINAC_REFINE) != 0 || mixed;
    if (refine && method == LINAC_METHOD_PLAIN)
        return fail(LINAC_ERR_INVALID, "LINAC_REFINE needs the blocked float method (its panel factors)");
    T* Aorig = nullptr;
    if (refine) {  // the unscaled input, for the residuals
        LINAC_TRY(ctx->get(SLOT_AUX, (size_t)rows * lda * sizeof(T), &p));
        Aorig = (T*)p;
        LINAC_HIP(hipMemcpy2DAsync(Aorig, lda * sizeof(T), A, lda * sizeof(T), cols * sizeof(T), rows,
                                   hipMemcpyDeviceToDevice, s));
    }
    if (prof) LINAC_TRY(prof->begin(s));
    k_row_scale<T><<<(unsigned)rows, 256, 0, s>>>(A, lda, rows, cols);
    k_float_init<<<(unsigned)std::min<int64_t>(1024, ceil_div(rows, 256)), 256, 0, s>>>(st, used, rows);
    // blocked: panels of 64 columns with a cooperative pivot kernel and an
    // FP64 matrix-core trailing update; the plain sweep stays the method for
    // small matrices, where per-column launches cost less than the panels
    bool blocked = method == LINAC_METHOD_BLOCKED || refine ||
                   (method == LINAC_METHOD_AUTO && rows >= tuning().float_blocked_min_rows &&
                    cols >= tuning().float_blocked_min_cols);
    FloatFactors fac;
    if (blocked) {
        const int rc = float_blocked_run<T>(A, lda, rows, cols, tol, s, &st->rank, used, piv_rows, piv_cols, ctx, prof,
                                            refine ? &fac : nullptr, mixed);
        if (rc < 0) return rc;
        if (rc == 1) {
            if (method == LINAC_METHOD_BLOCKED || refine)
                return fail(LINAC_ERR_INVALID, "blocked float method does not fit this device for %lld rows",
                            (long long)rows);
            blocked = false;
        }
    }
    const int64_t stride = std::max<int64_t>(1, cols / 256);
    for (int64_t j = 0; j < cols && !blocked; ++j) {
        k_float_select<T><<<1, 1024, 0, s>>>(A, lda, rows, j, tol, st, used, piv_rows, piv_cols);
        const int64_t w = cols - j - 1;
        const int log2G = w >= 1024 ? 8 : (w >= 256 ? 7 : 6);
        const bool sampled = prof && (j % stride) == 0;
        if (sampled)
            LINAC_TRY(prof->sample_begin(s, j, 2.0 * sizeof(T) * (double)(rows - 1) * (double)(cols - j),
                                         (double)(rows - 1) * (double)(cols - j)));
        k_float_elim<T><<<(unsigned)ceil_div(rows, 256 >> log2G), 256, 0, s>>>(A, lda, rows, cols, j, st, log2G);
        if (sampled) LINAC_TRY(prof->sample_end(s));
    }
    LINAC_HIP(hipGetLastError());
    if (prof) {
        LINAC_TRY(prof->finish(s));
        prof->result[5] = blocked ? 2 : 1;
    }
    if (refine) {
        int64_t rk = 0;
        LINAC_HIP(hipMemcpyAsync(&rk, &st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
        std::vector<int64_t> pcs(rk);
        if (rk) LINAC_HIP(hipMemcpy(pcs.data(), piv_cols, rk * 8, hipMemcpyDeviceToHost));
        LINAC_TRY(float_refine<T>(A, lda, Aorig, lda, rows, cols, piv_rows, piv_cols, pcs, fac,
                                  mixed ? MIXED_PASSES : REFINE_PASSES, s, ctx, mixed ? MIXED_STOP : 0.0));
        if (mixed && thread_refine_stats().last_correction > MIXED_STOP) {
            // the FP32 updates left the refinement short of FP64 accuracy
            // (cond(A) too close to 2^24): the elimination again from the
            // original matrix with FP64 updates, then the passes of LINAC_REFINE
            LINAC_HIP(hipMemcpy2DAsync(A, lda * sizeof(T), Aorig, lda * sizeof(T), cols * sizeof(T), rows,
                                       hipMemcpyDeviceToDevice, s));
            k_row_scale<T><<<(unsigned)rows, 256, 0, s>>>(A, lda, rows, cols);
            k_float_init<<<(unsigned)std::min<int64_t>(1024, ceil_div(rows, 256)), 256, 0, s>>>(st, used, rows);
            fac = FloatFactors{};
            const int rc = float_blocked_run<T>(A, lda, rows, cols, tol, s, &st->rank, used, piv_rows, piv_cols, ctx,
                                                nullptr, &fac, false);
            if (rc != LINAC_OK) return rc < 0 ? rc : fail(LINAC_ERR_DEVICE, "FP64 rerun of a mixed call failed");
            LINAC_HIP(hipMemcpyAsync(&rk, &st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
            LINAC_HIP(hipStreamSynchronize(s));
            pcs.assign((size_t)rk, 0);
            if (rk) LINAC_HIP(hipMemcpy(pcs.data(), piv_cols, rk * 8, hipMemcpyDeviceToHost));
            LINAC_TRY(float_refine<T>(A, lda, Aorig, lda, rows, cols, piv_rows, piv_cols, pcs, fac, REFINE_PASSES, s,
                                      ctx));
            thread_refine_stats().fallback = 1;
        }
    } else {
        thread_refine_stats() = RefineStats{};
    }
    void* outp = nullptr;
    LINAC_TRY(ctx->get(SLOT_OUT, (size_t)rows * lda * sizeof(T), &outp));
    k_float_finalize<T><<<(unsigned)rows, 256, 0, s>>>(A, lda, (T*)outp, rows, cols, st, piv_rows, piv_cols);
    LINAC_HIP(hipGetLastError());
    int64_t rank = 0;
    LINAC_HIP(hipMemcpyAsync(&rank, &st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> pc(rank);
    if (rank) LINAC_HIP(hipMemcpy(pc.data(), piv_cols, rank * 8, hipMemcpyDeviceToHost));
    if (prof && !blocked) {
        std::vector<char> is_piv(cols, 0);
        for (int64_t c : pc) is_piv[c] = 1;
        double bytes = 0, macs = 0;
        for (auto& x : prof->samples)
            if (is_piv[x.tag]) { bytes += x.bytes; macs += x.macs; }
        prof->result[3] = bytes;
        prof->result[4] = macs;
    }
    if (on_device) {
        LINAC_HIP(hipMemcpy2DAsync(a, ld * sizeof(T), outp, lda * sizeof(T), cols * sizeof(T), rows,
                                   hipMemcpyDeviceToDevice, s));
    } else {
        LINAC_HIP(hipMemcpy2DAsync(a, ld * sizeof(T), outp, lda * sizeof(T), cols * sizeof(T), rows,
                                   hipMemcpyDeviceToHost, s));
    }
    LINAC_HIP(hipStreamSynchronize(s));
    if (pivots_out)
        for (int64_t k = 0; k < rank; ++k) pivots_out[k] = pc[k];
    if (rank_out) *rank_out = rank;
    return LINAC_OK;
}

}  // namespace linac

// A blocked float RREF of a device matrix that runs a few panels per call, so
// that a caller can interleave a long solve with other work (the serving
// loop does, between batch steps).  Its workspace is the caller's, or, when
// the caller gives none, the context's (slots SLOT_JOB_A/B: then one such
// job per device at a time).
struct linac_float_job {
    linac::DeviceContext* ctx = nullptr;
    bool own_ws = false;  // caller's workspace
    bool cplx = false;
    void* a = nullptr;
    int64_t rows = 0, cols = 0, ld = 0;
    linac::FloatState* st = nullptr;
    int64_t* piv_rows = nullptr;
    int64_t* piv_cols = nullptr;
    uint8_t* used = nullptr;
    linac::FloatSweep<double> sd;
    linac::FloatSweep<linac::CplxF> sc;
};

namespace linac {

// Job state (rank, pivot lists, used flags) and the sweep's workspace, in
// that order, in one caller-supplied buffer.
static size_t job_state_bytes(int64_t rows, int64_t cols) {
    const int64_t cap = std::max<int64_t>(1, std::min(rows, cols));
    return 256 + 2 * round_up(cap * 8, 256) + round_up(rows, 256);
}

template <class T>
static int float_job_setup(linac_float_job* j, double tol, hipStream_t s, void* ws) {
    T* A = (T*)j->a;
    k_row_scale<T><<<(unsigned)j->rows, 256, 0, s>>>(A, j->ld, j->rows, j->cols);
    k_float_init<<<(unsigned)std::min<int64_t>(1024, ceil_div(j->rows, 256)), 256, 0, s>>>(j->st, j->used, j->rows);
    LINAC_HIP(hipGetLastError());
    FloatSweep<T>& sw = *reinterpret_cast<FloatSweep<T>*>(std::is_same<T, double>::value ? (void*)&j->sd : (void*)&j->sc);
    const int rc = sw.setup(A, j->ld, j->rows, j->cols, tol, s, &j->st->rank, j->used, j->piv_rows, j->piv_cols,
                            j->ctx, SLOT_JOB_A, nullptr, ws);
    if (rc == 1)
        return fail(LINAC_ERR_INVALID, "a float job of %lld rows does not fit this device's cooperative grid",
                    (long long)j->rows);
    return rc;
}

template <class T>
static int float_job_finish(linac_float_job* j, hipStream_t s, int64_t* pivots_out, int64_t* rank_out) {
    FloatSweep<T>& sw = *reinterpret_cast<FloatSweep<T>*>(std::is_same<T, double>::value ? (void*)&j->sd : (void*)&j->sc);
    LINAC_TRY(sw.finish(s, nullptr));
    void* outp = nullptr;
    LINAC_TRY(j->ctx->get(SLOT_OUT, (size_t)j->rows * j->ld * sizeof(T), &outp));
    k_float_finalize<T><<<(unsigned)j->rows, 256, 0, s>>>((const T*)j->a, j->ld, (T*)outp, j->rows, j->cols, j->st,
                                                          j->piv_rows, j->piv_cols);
    LINAC_HIP(hipGetLastError());
    LINAC_HIP(hipMemcpy2DAsync(j->a, j->ld * sizeof(T), outp, j->ld * sizeof(T), j->cols * sizeof(T), j->rows,
                               hipMemcpyDeviceToDevice, s));
    int64_t rank = 0;
    LINAC_HIP(hipMemcpyAsync(j->ctx->pinned, &j->st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipStreamSynchronize(s));
    rank = j->ctx->pinned[0];
    if (pivots_out && rank) LINAC_HIP(hipMemcpy(pivots_out, j->piv_cols, rank * 8, hipMemcpyDeviceToHost));
    if (rank_out) *rank_out = rank;
    return LINAC_OK;
}

// Open jobs: linac_shutdown releases them with the contexts; a handle used
// after that is recognised as released (destroy is then a no-op).
static std::mutex g_jobs_mu;
static std::vector<linac_float_job*> g_jobs;

static bool job_open(linac_float_job* j) {
    std::lock_guard<std::mutex> lk(g_jobs_mu);
    return std::find(g_jobs.begin(), g_jobs.end(), j) != g_jobs.end();
}

static bool job_forget(linac_float_job* j) {
    std::lock_guard<std::mutex> lk(g_jobs_mu);
    auto it = std::find(g_jobs.begin(), g_jobs.end(), j);
    if (it == g_jobs.end()) return false;
    g_jobs.erase(it);
    return true;
}

static void free_float_jobs(DeviceContext* ctx) {
    std::lock_guard<std::mutex> lk(g_jobs_mu);
    for (auto it = g_jobs.begin(); it != g_jobs.end();) {
        if ((*it)->ctx == ctx) {
            delete *it;
            it = g_jobs.erase(it);
        } else {
            ++it;
        }
    }
    ctx->float_job = nullptr;
}

}  // namespace linac

using namespace linac;

extern "C" {

int linac_float_job_workspace(int64_t rows, int64_t cols, int is_complex, int64_t* bytes_out) {
    if (!bytes_out) return fail(LINAC_ERR_INVALID, "bytes_out is NULL");
    if (rows <= 0 || cols <= 0) return fail(LINAC_ERR_INVALID, "empty matrix");
    const size_t sweep = is_complex ? FloatSweep<CplxF>::workspace_bytes(rows, cols)
                                    : FloatSweep<double>::workspace_bytes(rows, cols);
    *bytes_out = (int64_t)(job_state_bytes(rows, cols) + sweep);
    return LINAC_OK;
}

int linac_float_job_create(linac_float_job** out, double* a, int64_t rows, int64_t cols, int64_t ld, int is_complex,
                           double tol, uint32_t flags, void* workspace, int64_t workspace_bytes) {
    if (!out) return fail(LINAC_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!(flags & LINAC_DEVICE_PTR)) return fail(LINAC_ERR_INVALID, "a float job works on a device matrix (LINAC_DEVICE_PTR)");
    if (rows <= 0 || cols <= 0 || !a) return fail(LINAC_ERR_INVALID, "empty matrix or NULL pointer");
    if (ld < cols) return fail(LINAC_ERR_INVALID, "ld (%lld) < cols (%lld)", (long long)ld, (long long)cols);
    if (!(tol >= 0.0)) return fail(LINAC_ERR_INVALID, "tolerance must be >= 0");
    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (workspace) {
        int64_t need = 0;
        LINAC_TRY(linac_float_job_workspace(rows, cols, is_complex, &need));
        if (workspace_bytes < need)
            return fail(LINAC_ERR_INVALID, "workspace of %lld bytes, %lld needed", (long long)workspace_bytes,
                        (long long)need);
        if ((uintptr_t)workspace % 256)
            return fail(LINAC_ERR_INVALID, "workspace must be 256-byte aligned");
    } else if (ctx->float_job) {
        return fail(LINAC_ERR_INVALID, "a float job without its own workspace is already open on this device");
    }
    auto* j = new linac_float_job();
    j->ctx = ctx;
    j->own_ws = workspace != nullptr;
    j->cplx = is_complex != 0;
    j->a = a;
    j->rows = rows;
    j->cols = cols;
    j->ld = ld;
    const int64_t cap = std::max<int64_t>(1, std::min(rows, cols));
    void* p = workspace;
    void* sweep_ws = workspace ? (char*)workspace + job_state_bytes(rows, cols) : nullptr;
    int rc = p ? LINAC_OK : ctx->get(SLOT_JOB_B, job_state_bytes(rows, cols), &p);
    if (rc == LINAC_OK) {
        j->st = (FloatState*)p;
        j->piv_rows = (int64_t*)((char*)p + 256);
        j->piv_cols = (int64_t*)((char*)p + 256 + round_up(cap * 8, 256));
        j->used = (uint8_t*)((char*)p + 256 + 2 * round_up(cap * 8, 256));
        rc = j->cplx ? float_job_setup<CplxF>(j, tol, ctx->stream, sweep_ws)
                     : float_job_setup<double>(j, tol, ctx->stream, sweep_ws);
    }
    if (rc != LINAC_OK) {
        delete j;
        return rc;
    }
    add_shutdown_hook(free_float_jobs);
    if (!j->own_ws) ctx->float_job = j;
    {
        std::lock_guard<std::mutex> g(g_jobs_mu);
        g_jobs.push_back(j);
    }
    *out = j;
    return LINAC_OK;
}

int linac_float_job_step(linac_float_job* j, int64_t max_panels, int64_t* remaining_out) {
    if (!j || !job_open(j)) return fail(LINAC_ERR_INVALID, "job is NULL or already released");
    if (max_panels <= 0) return fail(LINAC_ERR_INVALID, "max_panels must be positive");
    std::lock_guard<std::recursive_mutex> lk(j->ctx->mu);
    hipStream_t s = j->ctx->stream;
    if (j->cplx) LINAC_TRY(j->sc.run(s, max_panels, nullptr));
    else LINAC_TRY(j->sd.run(s, max_panels, nullptr));
    LINAC_HIP(hipStreamSynchronize(s));
    if (remaining_out) *remaining_out = j->cplx ? j->sc.npanels - j->sc.next : j->sd.npanels - j->sd.next;
    return LINAC_OK;
}

int linac_float_job_finish(linac_float_job* j, int64_t* pivots_out, int64_t* rank_out) {
    if (!j || !job_forget(j)) return fail(LINAC_ERR_INVALID, "job is NULL or already released");
    std::lock_guard<std::recursive_mutex> lk(j->ctx->mu);
    hipStream_t s = j->ctx->stream;
    if (!(j->cplx ? j->sc.done() : j->sd.done())) {
        if (j->cplx) LINAC_TRY(j->sc.run(s, -1, nullptr));
        else LINAC_TRY(j->sd.run(s, -1, nullptr));
    }
    const int rc = j->cplx ? float_job_finish<CplxF>(j, s, pivots_out, rank_out)
                           : float_job_finish<double>(j, s, pivots_out, rank_out);
    if (j->ctx->float_job == j) j->ctx->float_job = nullptr;
    delete j;
    return rc;
}

int linac_float_job_destroy(linac_float_job* j) {
    if (!j || !job_forget(j)) return LINAC_OK;  // NULL, finished, or released by linac_shutdown
    std::lock_guard<std::recursive_mutex> lk(j->ctx->mu);
    (void)hipStreamSynchronize(j->ctx->stream);
    if (j->ctx->float_job == j) j->ctx->float_job = nullptr;
    delete j;
    return LINAC_OK;
}

int linac_rref_f64(double* a, int64_t rows, int64_t cols, int64_t ld, double tol, uint32_t flags, void* stream,
                   int64_t* pivots_out, int64_t* rank_out) {
    return rref_float_entry<double>(a, rows, cols, ld, tol, flags, stream, pivots_out, rank_out);
}

int linac_rref_c128(double* a, int64_t rows, int64_t cols, int64_t ld, double tol, uint32_t flags, void* stream,
                    int64_t* pivots_out, int64_t* rank_out) {
    return rref_float_entry<CplxF>((CplxF*)a, rows, cols, ld, tol, flags, stream, pivots_out, rank_out);
}

}  // extern "C"
