// Gauss-Jordan RREF over GF(p): the unblocked ("plain") method, the shared
// finalisation, and the C entry points linac_rref_u32 / linac_rref_u64.
//
// Replaces linac/pycuda_row_reduce.py:18-151 and the GF(p) kernels of
// linac/row_reduce.cu (ThreadsReduceToMaxIndex :287-342, SwitchRows :273-285,
// CompareHeadToTolerance :404-418, RescaleRow :117-136, RowReduce :138-196,
// IncrementCounters :395-402).  Design differences:
//
//  * no row swaps: a pivot row stays where it is, is flagged `used`, and its
//    index is appended to the pivot list; the final gather writes row k of the
//    result from the k-th pivot row (rows that never pivot end all-zero);
//  * pivot rows are not normalised during the sweep: the other rows are
//    updated with the factor a/h (h the head), and each pivot row is scaled by
//    h^{-1} once, in the gather.  A pivot row's head column is never touched
//    again after its step, so h is still there at the end;
//  * the pivot search covers every unused row (the reference searches a
//    2048-row window, row_reduce.cu:300-303); since the RREF over a field is
//    unique, choosing the first non-zero instead of the largest value yields
//    the same matrix;
//  * one select + one eliminate launch per column, prime passed at run time.
#include <algorithm>
#include <vector>

#include "field.hpp"
#include "gf_common.hpp"
#include "runtime.hpp"

namespace linac {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Every entry reduced into [0, p).  Rows along blockIdx.y, 16-byte vectors
// along x where rows are 16-byte aligned (vec); a vector is written back only
// when one of its entries changed, so a canonical matrix costs one read.
template <class F>
__global__ __launch_bounds__(256) void k_reduce_mod(typename F::T* __restrict__ A, int64_t ld, int64_t rows,
                                                    int64_t cols, F f, int vec) {
    using T = typename F::T;
    constexpr int V = 16 / (int)sizeof(T);
    const int64_t nv = vec ? cols / V : 0;
    constexpr int U = 4;  // vectors in flight per thread
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) {
        T* row = A + r * ld;
        for (int64_t v0 = (int64_t)blockIdx.x * 256 + threadIdx.x; v0 < nv; v0 += U * stride) {
            Vec16<T> x[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (v0 + u * stride < nv) x[u] = reinterpret_cast<const Vec16<T>*>(row)[v0 + u * stride];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (v0 + u * stride >= nv) break;
                bool changed = false;
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    const T y = f.reduce(x[u].v[e]);
                    changed = changed || y != x[u].v[e];
                    x[u].v[e] = y;
                }
                if (changed) reinterpret_cast<Vec16<T>*>(row)[v0 + u * stride] = x[u];
            }
        }
        for (int64_t c = nv * V + (int64_t)blockIdx.x * 256 + threadIdx.x; c < cols; c += (int64_t)gridDim.x * 256) {
            const T y = f.reduce(row[c]);
            if (y != row[c]) row[c] = y;
        }
    }
}

__global__ __launch_bounds__(256) void k_init_state(ElimState* st, uint8_t* used, int64_t N) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x == 0) {
        st->prow = -1;
        st->rank = 0;
        st->first_unused = 0;
        st->factor = 0;
        for (int i = 0; i < 4; ++i) st->aux[i] = 0;
    }
    for (int64_t r = x; r < N; r += (int64_t)gridDim.x * blockDim.x) used[r] = 0;
}

// Pivot of column j: the smallest unused row with a non-zero entry.  Rows are
// probed 1024 at a time from the first unused row upward; each wave ballots
// its hits and the first hit of the first non-empty chunk wins.
template <class F>
__global__ __launch_bounds__(1024) void k_select(const typename F::T* __restrict__ A, int64_t ld, int64_t N,
                                                 int64_t j, F f, ElimState* st, uint8_t* used,
                                                 int64_t* piv_rows, int64_t* piv_cols) {
    __shared__ unsigned long long s_best;
    const int tid = threadIdx.x;
    if (tid == 0) s_best = ~0ull;
    __syncthreads();
    const int64_t start = st->first_unused;
    for (int64_t base = start; base < N; base += 1024) {
        const int64_t r = base + tid;
        bool hit = false;
        if (r < N && !used[r]) hit = A[r * ld + j] != 0;
        const unsigned long long mask = __ballot(hit);
        if ((tid & 63) == 0 && mask)
            atomicMin(&s_best, (unsigned long long)(base + (tid & ~63) + __ffsll((long long)mask) - 1));
        __syncthreads();
        const bool done = s_best != ~0ull;
        __syncthreads();
        if (done) break;
    }
    if (tid == 0) {
        if (s_best != ~0ull) {
            const int64_t r = (int64_t)s_best;
            st->factor = (uint64_t)f.factor_of_head(A[r * ld + j]);
            st->prow = r;
            used[r] = 1;
            const int64_t k = st->rank;
            piv_rows[k] = r;
            piv_cols[k] = j;
            st->rank = k + 1;
            if (r == start) {
                int64_t q = r + 1;
                while (q < N && used[q]) ++q;
                st->first_unused = q;
            }
        } else {
            st->prow = -1;
        }
    }
}

// Row update of column j: A[r, c] -= (A[r, j] / h) * A[prow, c] for c > j and
// A[r, j] = 0, every row r != prow.  G = 2^log2G threads share a row; each
// thread streams 16-byte vectors (4 x u32 or 2 x u64).
template <class F>
__global__ __launch_bounds__(256) void k_elim(typename F::T* __restrict__ A, int64_t ld, int64_t N, int64_t M,
                                              int64_t j, F f, const ElimState* __restrict__ st, int log2G) {
    using T = typename F::T;
    constexpr int V = 16 / sizeof(T);
    using VT = Vec16<T>;
    const int64_t prow = st->prow;
    if (prow < 0) return;
    const int G = 1 << log2G;
    const int64_t r = ((int64_t)blockIdx.x << (8 - log2G)) + (threadIdx.x >> log2G);
    const int t = threadIdx.x & (G - 1);
    if (r >= N || r == prow) return;
    T* row = A + r * ld;
    const T a = row[j];
    if (a == 0) return;
    const T fm = f.mont(a, (T)st->factor);
    const T* pr = A + prow * ld;
    const int64_t v0 = (j + 1) / V, v1 = (M + V - 1) / V;
    for (int64_t v = v0 + t; v < v1; v += G) {
        VT x = reinterpret_cast<const VT*>(row)[v];
        const VT y = reinterpret_cast<const VT*>(pr)[v];
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const int64_t c = v * V + e;
            const T nv = f.sub(x.v[e], f.mont(fm, y.v[e]));
            x.v[e] = (c > j && c < M) ? nv : x.v[e];
        }
        reinterpret_cast<VT*>(row)[v] = x;
    }
    if (t == 0) row[j] = 0;
}

// Row k of the result = (k-th pivot row) * inverse(head) when `normalize`,
// or the k-th pivot row as is; rows >= rank are zero.  In place when
// out == A and the k-th pivot row is row k.
template <class F>
__global__ __launch_bounds__(256) void k_finalize(const typename F::T* A, int64_t ld, typename F::T* out,
                                                  int64_t ld_out, int64_t N, int64_t M, F f,
                                                  const ElimState* __restrict__ st,
                                                  const int64_t* __restrict__ piv_rows,
                                                  const int64_t* __restrict__ piv_cols, int normalize,
                                                  int vec, const int* gate = nullptr, int want = 0) {
    using T = typename F::T;
    constexpr int V = 16 / sizeof(T);
    using VT = Vec16<T>;
    __shared__ T s_scale;
    if (gate && *gate != want) return;  // the other form of the gather applies
    const int64_t k = blockIdx.x;
    const int64_t rank = st->rank;
    // vec: rows are 16-byte aligned and padded to whole vectors (ld % V == 0),
    // so a row may be written as round_up(M, V) elements; otherwise per element
    const int64_t nv = (M + V - 1) / V;
    T* orow = out + k * ld_out;
    if (k >= rank) {
        if (vec) {
            VT z;
#pragma unroll
            for (int e = 0; e < V; ++e) z.v[e] = 0;
            for (int64_t v = threadIdx.x; v < nv; v += blockDim.x) reinterpret_cast<VT*>(orow)[v] = z;
        } else {
            for (int64_t c = threadIdx.x; c < M; c += blockDim.x) orow[c] = 0;
        }
        return;
    }
    const T* s = A + piv_rows[k] * ld;
    if (normalize) {
        if (threadIdx.x == 0) s_scale = f.to_mont(f.inverse(s[piv_cols[k]]));
        __syncthreads();
    }
    const T sc = normalize ? s_scale : (T)0;
    if (vec) {
        for (int64_t v = threadIdx.x; v < nv; v += blockDim.x) {
            VT x = reinterpret_cast<const VT*>(s)[v];
            if (normalize) {
#pragma unroll
                for (int e = 0; e < V; ++e) x.v[e] = f.mont(sc, x.v[e]);
            }
            reinterpret_cast<VT*>(orow)[v] = x;
        }
    } else {
        for (int64_t c = threadIdx.x; c < M; c += blockDim.x) orow[c] = normalize ? f.mont(sc, s[c]) : s[c];
    }
}

// *ident = 1 when the k-th pivot row is row k for every k < rank (the gather
// can run in place), else 0.  One workgroup.
__global__ __launch_bounds__(256) void k_identity_flag(const ElimState* __restrict__ st,
                                                       const int64_t* __restrict__ piv_rows, int* ident) {
    __shared__ int s_ok;
    if (threadIdx.x == 0) s_ok = 1;
    __syncthreads();
    const int64_t rank = st->rank;
    bool ok = true;
    for (int64_t k = threadIdx.x; k < rank; k += blockDim.x) ok = ok && piv_rows[k] == k;
    if (!ok) s_ok = 0;
    __syncthreads();
    if (threadIdx.x == 0) *ident = s_ok;
}

// dst[r, :cols] = src[r, :cols] unless *gate != want.
template <class T>
__global__ __launch_bounds__(256) void k_copy_rows_gated(T* __restrict__ dst, int64_t ld_dst, const T* __restrict__ src,
                                                         int64_t ld_src, int64_t rows, int64_t cols,
                                                         const int* __restrict__ gate, int want) {
    if (*gate != want) return;
    const int64_t total = rows * cols;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = x / cols, c = x - r * cols;
        dst[r * ld_dst + c] = src[r * ld_src + c];
    }
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------

template <class F>
int plain_launch_init(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                      const ElimWorkspace& ws) {
    using T = typename F::T;
    constexpr int V = 16 / (int)sizeof(T);
    const int vec = ((ld * (int64_t)sizeof(T)) % 16 == 0 && ((uintptr_t)A & 15) == 0) ? 1 : 0;
    const int64_t bx = std::min<int64_t>(16, std::max<int64_t>(1, ceil_div(vec ? ceil_div(M, V) : M, 1024)));
    const int64_t by = std::min<int64_t>(std::max<int64_t>(N, 1), std::max<int64_t>(1, 16384 / bx));
    k_reduce_mod<F><<<dim3((unsigned)bx, (unsigned)by), 256, 0, s>>>(A, ld, N, M, f, vec);
    k_init_state<<<(int)std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div(N, 256))), 256, 0, s>>>(ws.st, ws.used, N);
    LINAC_HIP(hipGetLastError());
    return LINAC_OK;
}

template <class F>
int plain_run(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
              const ElimWorkspace& ws, DeviceContext* ctx, Profiler* prof) {
    using T = typename F::T;
    constexpr int V = 16 / sizeof(T);
    // sample ~256 row-update launches spread over the sweep
    const int64_t stride = std::max<int64_t>(1, M / 256);
    hipEvent_t probe = ctx->probe;
    bool probe_pending = false;
    ctx->pinned[0] = 0;
    for (int64_t j = 0; j < M; ++j) {
        k_select<F><<<1, 1024, 0, s>>>(A, ld, N, j, f, ws.st, ws.used, ws.piv_rows, ws.piv_cols);
        const int64_t nvec = ceil_div(M, V) - (j + 1) / V;
        const int log2G = nvec >= 1024 ? 8 : (nvec >= 256 ? 7 : 6);
        const int64_t rows_per_block = 256 >> log2G;
        const bool sampled = prof && (j % stride) == 0;
        if (sampled) {
            // algorithmic traffic of a pivot step: read + write the trailing part
            // of the N-1 other rows (2 * bpe * (N-1) * (M-j)); no pivot -> 0,
            // corrected after the sweep from the pivot list.
            LINAC_TRY(prof->sample_begin(s, j, 2.0 * sizeof(T) * (double)(N - 1) * (double)(M - j),
                                         (double)(N - 1) * (double)(M - j)));
        }
        k_elim<F><<<(unsigned)ceil_div(N, rows_per_block), 256, 0, s>>>(A, ld, N, M, j, f, ws.st, log2G);
        if (sampled) LINAC_TRY(prof->sample_end(s));
        // Once every row is a pivot row the remaining columns cannot change;
        // poll the rank without blocking the launch stream (every 32 columns
        // for at most 64 rows, else every 256).
        const int64_t poll = N <= 64 ? 31 : 255;
        if (M > N && (j & poll) == poll) {
            if (probe_pending && hipEventQuery(probe) == hipSuccess) {
                probe_pending = false;
                if (ctx->pinned[0] >= N) break;
            }
            if (!probe_pending) {
                LINAC_HIP(hipMemcpyAsync(ctx->pinned, &ws.st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
                LINAC_HIP(hipEventRecord(probe, s));
                probe_pending = true;
            }
        }
    }
    LINAC_HIP(hipGetLastError());
    LINAC_HIP(hipStreamSynchronize(s));
    return LINAC_OK;
}

// Enqueues the final gather on s (no synchronisation): rows/cols are the
// pivot list in the order the result rows take (the device copy of the list
// is rewritten when `resorted`).
template <class F>
int gf_finalize_apply(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                      const ElimWorkspace& ws, DeviceContext* ctx, bool normalize, const std::vector<int64_t>& rows,
                      const std::vector<int64_t>& cols, bool resorted) {
    using T = typename F::T;
    const int64_t rank = (int64_t)rows.size();
    if (resorted && rank > 0) {
        LINAC_HIP(hipMemcpyAsync(ws.piv_rows, rows.data(), rank * 8, hipMemcpyHostToDevice, s));
        LINAC_HIP(hipMemcpyAsync(ws.piv_cols, cols.data(), rank * 8, hipMemcpyHostToDevice, s));
    }
    bool identity = true;
    for (int64_t k = 0; k < rank && identity; ++k) identity = rows[k] == k;
    constexpr int V = 16 / sizeof(T);
    const int vec = (ld % V == 0 && ((uintptr_t)A & 15) == 0) ? 1 : 0;
    if (identity) {
        if (normalize || rank < N)
            k_finalize<F><<<(unsigned)N, 256, 0, s>>>(A, ld, A, ld, N, M, f, ws.st, ws.piv_rows, ws.piv_cols,
                                                       normalize ? 1 : 0, vec);
    } else {
        void* outp = nullptr;
        LINAC_TRY(ctx->get(SLOT_OUT, (size_t)N * ld * sizeof(T), &outp));
        T* out = (T*)outp;
        k_finalize<F><<<(unsigned)N, 256, 0, s>>>(A, ld, out, ld, N, M, f, ws.st, ws.piv_rows, ws.piv_cols,
                                                   normalize ? 1 : 0, vec);
        LINAC_HIP(hipMemcpyAsync(A, out, (size_t)N * ld * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    LINAC_HIP(hipGetLastError());
    return LINAC_OK;
}

// Sorts a pivot list by column in place; true when the order changed.
bool sort_pivots_by_column(std::vector<int64_t>& rows, std::vector<int64_t>& cols) {
    const int64_t rank = (int64_t)rows.size();
    bool sorted = true;
    for (int64_t k = 1; k < rank && sorted; ++k) sorted = cols[k - 1] < cols[k];
    if (sorted) return false;
    std::vector<int64_t> order(rank);
    for (int64_t k = 0; k < rank; ++k) order[k] = k;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return cols[a] < cols[b]; });
    std::vector<int64_t> r2(rank), c2(rank);
    for (int64_t k = 0; k < rank; ++k) {
        r2[k] = rows[order[k]];
        c2[k] = cols[order[k]];
    }
    rows.swap(r2);
    cols.swap(c2);
    return true;
}

template <class F>
int gf_finalize(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                const ElimWorkspace& ws, DeviceContext* ctx, bool normalize, bool sort_by_column,
                int64_t* rank_host, std::vector<int64_t>* piv_cols_host, const std::vector<EchelonPanel>* echelon) {
    LINAC_HIP(hipMemcpyAsync(ctx->pinned, &ws.st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipStreamSynchronize(s));
    const int64_t rank = ctx->pinned[0];
    std::vector<int64_t> rows(rank), cols(rank);
    if (rank > 0) {
        LINAC_HIP(hipMemcpyAsync(rows.data(), ws.piv_rows, rank * 8, hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipMemcpyAsync(cols.data(), ws.piv_cols, rank * 8, hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
    }
    const bool resorted = sort_by_column && sort_pivots_by_column(rows, cols);
    if (echelon && !echelon->empty())
        LINAC_TRY(echelon_finalize(A, ld, N, M, f, s, ws, ctx, rows, cols, resorted, *echelon));
    else
        LINAC_TRY(gf_finalize_apply(A, ld, N, M, f, s, ws, ctx, normalize, rows, cols, resorted));
    LINAC_HIP(hipStreamSynchronize(s));
    *rank_host = rank;
    *piv_cols_host = cols;
    return LINAC_OK;
}

template int gf_finalize_apply<FieldU32>(uint32_t*, int64_t, int64_t, int64_t, const FieldU32&, hipStream_t,
                                         const ElimWorkspace&, DeviceContext*, bool, const std::vector<int64_t>&,
                                         const std::vector<int64_t>&, bool);
template int gf_finalize<FieldU32>(uint32_t*, int64_t, int64_t, int64_t, const FieldU32&, hipStream_t,
                                   const ElimWorkspace&, DeviceContext*, bool, bool, int64_t*, std::vector<int64_t>*,
                                   const std::vector<EchelonPanel>*);
template int gf_finalize_apply<FieldU64>(uint64_t*, int64_t, int64_t, int64_t, const FieldU64&, hipStream_t,
                                         const ElimWorkspace&, DeviceContext*, bool, const std::vector<int64_t>&,
                                         const std::vector<int64_t>&, bool);
template int gf_finalize<FieldU64>(uint64_t*, int64_t, int64_t, int64_t, const FieldU64&, hipStream_t,
                                   const ElimWorkspace&, DeviceContext*, bool, bool, int64_t*, std::vector<int64_t>*,
                                   const std::vector<EchelonPanel>*);

// Shared entry: validation, host<->device staging, method dispatch.
template <class F>
static int rref_gf_entry(typename F::T* a, int64_t rows, int64_t cols, int64_t ld, const F& f, uint32_t flags,
                         void* stream_arg, int64_t* pivots_out, int64_t* rank_out) {
    using T = typename F::T;
    if (rows < 0 || cols < 0) return fail(LINAC_ERR_INVALID, "negative shape (%lld, %lld)", (long long)rows, (long long)cols);
    if (ld < cols) return fail(LINAC_ERR_INVALID, "ld (%lld) < cols (%lld)", (long long)ld, (long long)cols);
    if (rows > 0 && cols > 0 && !a) return fail(LINAC_ERR_INVALID, "matrix pointer is NULL");
    if (flags & LINAC_REFINE) return fail(LINAC_ERR_INVALID, "LINAC_REFINE applies to float64 / complex128 only");
    if (rank_out) *rank_out = 0;
    if (rows == 0 || cols == 0) return LINAC_OK;
    const bool on_device = (flags & LINAC_DEVICE_PTR) != 0;
    if (on_device && ((ld * (int64_t)sizeof(T)) % 16 != 0 || ((uintptr_t)a & 15) != 0))
        return fail(LINAC_ERR_INVALID, "device matrix rows must be 16-byte aligned (ld*%d %% 16 == 0)", (int)sizeof(T));
    const uint32_t method = flags & LINAC_METHOD_MASK;
    if (method == LINAC_METHOD_MIXED) return fail(LINAC_ERR_INVALID, "the mixed method is for float64 matrices");
    if (method != LINAC_METHOD_PLAIN && method != LINAC_METHOD_AUTO && method != LINAC_METHOD_BLOCKED)
        return fail(LINAC_ERR_INVALID, "unknown method flags 0x%x", method);
    Profiler* prof = (flags & LINAC_PROFILE) ? &thread_profiler() : nullptr;

    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipStream_t s = stream_arg ? (hipStream_t)stream_arg : ctx->stream;

    T* A = a;
    int64_t lda = ld;
    if (!on_device) {
        lda = round_up(cols, 64);
        void* dev = nullptr;
        LINAC_TRY(ctx->get(SLOT_MATRIX, (size_t)rows * lda * sizeof(T), &dev));
        A = (T*)dev;
        LINAC_HIP(hipMemcpy2DAsync(A, lda * sizeof(T), a, ld * sizeof(T), cols * sizeof(T), rows,
                                   hipMemcpyHostToDevice, s));
    }
    ElimWorkspace ws;
    LINAC_TRY(make_workspace(ctx, rows, cols, &ws));
    if (prof) LINAC_TRY(prof->begin(s));
    LINAC_TRY(plain_launch_init(A, lda, rows, cols, f, s, ws));

    bool normalize = true, sorted = false;
    int rc = NOT_APPLICABLE;
    // (left unfinalised, the rows must be final: Gauss-Jordan form)
    std::vector<EchelonPanel> echelon;
    if (method == LINAC_METHOD_BLOCKED || method == LINAC_METHOD_AUTO)
        rc = blocked_dispatch(A, lda, rows, cols, f, s, ws, ctx, method == LINAC_METHOD_BLOCKED, prof,
                              (flags & LINAC_NO_FINALIZE) ? nullptr : &echelon);
    if (rc < 0) return rc;
    if (rc == LINAC_OK) {
        normalize = false;
        sorted = true;
    } else {
        if (method == LINAC_METHOD_BLOCKED) return fail(LINAC_ERR_INVALID, "blocked method does not apply here");
        if (method != LINAC_METHOD_PLAIN && method != LINAC_METHOD_AUTO)
            return fail(LINAC_ERR_INVALID, "unknown method flags 0x%x", method);
        LINAC_TRY(plain_run(A, lda, rows, cols, f, s, ws, ctx, prof));
    }
    if (prof) {
        LINAC_TRY(prof->finish(s));
        prof->result[5] = rc == LINAC_OK ? 2 : 1;
    }
    int64_t rank = 0;
    std::vector<int64_t> pc;
    if (!(flags & LINAC_NO_FINALIZE)) {
        LINAC_TRY(gf_finalize(A, lda, rows, cols, f, s, ws, ctx, normalize, sorted, &rank, &pc, &echelon));
    } else {
        LINAC_HIP(hipMemcpyAsync(ctx->pinned, &ws.st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
        rank = ctx->pinned[0];
        pc.resize(rank);
        if (rank) LINAC_HIP(hipMemcpy(pc.data(), ws.piv_cols, rank * 8, hipMemcpyDeviceToHost));
    }
    if (!on_device) {
        LINAC_HIP(hipMemcpy2DAsync(a, ld * sizeof(T), A, lda * sizeof(T), cols * sizeof(T), rows,
                                   hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
    }
    if (prof && rc != LINAC_OK) {
        // plain sweep: a sampled column without a pivot launched no work
        std::vector<char> is_piv(cols, 0);
        for (int64_t c : pc) is_piv[c] = 1;
        double bytes = 0, macs = 0;
        for (auto& x : prof->samples)
            if (is_piv[x.tag]) { bytes += x.bytes; macs += x.macs; }
        prof->result[3] = bytes;
        prof->result[4] = macs;
    }
    if (pivots_out)
        for (int64_t k = 0; k < rank; ++k) pivots_out[k] = pc[k];
    if (rank_out) *rank_out = rank;
    return LINAC_OK;
}

// Bytes of caller-provided scratch the plain method needs without any other
// allocation: elimination state, pivot lists, row flags, a gate word and a
// gather buffer of the whole matrix (used when pivot rows are not rows
// 0..rank-1).
size_t noalloc_scratch_bytes(int64_t rows, int64_t cols, int64_t ld, size_t elem) {
    const int64_t cap = std::max<int64_t>(1, std::min(rows, cols));
    return 256 + 2 * (size_t)round_up(cap * 8, 256) + (size_t)round_up(rows, 256) + 256 +
           (size_t)round_up(rows * ld * (int64_t)elem, 256);
}

// The plain method on a device matrix with every buffer carved from `scratch`
// (device memory of at least noalloc_scratch_bytes): no allocation, no
// context scratch.  Uses the context's stream, pinned status page and probe
// event, which exist once the context does.
template <class F>
static int rref_gf_noalloc(typename F::T* A, int64_t rows, int64_t cols, int64_t ld, const F& f, uint32_t flags,
                           hipStream_t s, DeviceContext* ctx, uint8_t* scratch, int64_t* pivots_out,
                           int64_t* rank_out) {
    using T = typename F::T;
    const int64_t cap = std::max<int64_t>(1, std::min(rows, cols));
    char* b = (char*)scratch;
    ElimWorkspace ws;
    ws.st = (ElimState*)b;
    b += 256;
    ws.piv_rows = (int64_t*)b;
    b += round_up(cap * 8, 256);
    ws.piv_cols = (int64_t*)b;
    b += round_up(cap * 8, 256);
    ws.used = (uint8_t*)b;
    b += round_up(rows, 256);
    ws.cap = cap;
    int* ident = (int*)b;
    b += 256;
    T* gather = (T*)b;
    LINAC_TRY(plain_launch_init(A, ld, rows, cols, f, s, ws));
    LINAC_TRY(plain_run(A, ld, rows, cols, f, s, ws, ctx, nullptr));
    if (!(flags & LINAC_NO_FINALIZE)) {
        constexpr int V = 16 / sizeof(T);
        const int vec = (ld % V == 0 && ((uintptr_t)A & 15) == 0) ? 1 : 0;
        k_identity_flag<<<1, 256, 0, s>>>(ws.st, ws.piv_rows, ident);
        // in place when the pivot rows are 0..rank-1, else through the gather buffer
        k_finalize<F><<<(unsigned)rows, 256, 0, s>>>(A, ld, A, ld, rows, cols, f, ws.st, ws.piv_rows, ws.piv_cols, 1,
                                                      vec, ident, 1);
        k_finalize<F><<<(unsigned)rows, 256, 0, s>>>(A, ld, gather, ld, rows, cols, f, ws.st, ws.piv_rows,
                                                      ws.piv_cols, 1, vec, ident, 0);
        const unsigned grid = (unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, ceil_div(rows * cols, 256)));
        k_copy_rows_gated<T><<<grid, 256, 0, s>>>(A, ld, gather, ld, rows, cols, ident, 0);
        LINAC_HIP(hipGetLastError());
    }
    LINAC_HIP(hipMemcpyAsync(ctx->pinned, &ws.st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipStreamSynchronize(s));
    const int64_t rank = ctx->pinned[0];
    if (pivots_out && rank > 0) {
        LINAC_HIP(hipMemcpyAsync(pivots_out, ws.piv_cols, rank * 8, hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
    }
    if (rank_out) *rank_out = rank;
    return LINAC_OK;
}

int rref_u32_noalloc(uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p, uint32_t flags, hipStream_t s,
                     DeviceContext* ctx, uint8_t* scratch, int64_t* pivots_out, int64_t* rank_out) {
    if (p == 2)
        return rref_gf_noalloc(a, rows, cols, ld, FieldGF2{2u}, flags, s, ctx, scratch, pivots_out, rank_out);
    return rref_gf_noalloc(a, rows, cols, ld, make_field_u32(p), flags, s, ctx, scratch, pivots_out, rank_out);
}

}  // namespace linac

using namespace linac;

extern "C" {

int linac_rref_u32(uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p, uint32_t flags, void* stream,
                   int64_t* pivots_out, int64_t* rank_out) {
    if (!is_prime_u64(p)) return fail(LINAC_ERR_INVALID, "field characteristic %u is not prime", p);
    if (p == 2) return rref_gf_entry(a, rows, cols, ld, FieldGF2{2u}, flags, stream, pivots_out, rank_out);
    return rref_gf_entry(a, rows, cols, ld, make_field_u32(p), flags, stream, pivots_out, rank_out);
}

int linac_rref_u64(uint64_t* a, int64_t rows, int64_t cols, int64_t ld, uint64_t p, uint32_t flags, void* stream,
                   int64_t* pivots_out, int64_t* rank_out) {
    if (p >= (1ull << 63)) return fail(LINAC_ERR_INVALID, "field characteristic %llu >= 2^63 is not supported", (unsigned long long)p);
    if (!is_prime_u64(p)) return fail(LINAC_ERR_INVALID, "field characteristic %llu is not prime", (unsigned long long)p);
    if (p == 2) return fail(LINAC_ERR_INVALID, "use linac_rref_u32 for p = 2");
    return rref_gf_entry(a, rows, cols, ld, make_field_u64(p), flags, stream, pivots_out, rank_out);
}

}  // extern "C"
