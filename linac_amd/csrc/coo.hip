// Coordinate-list (COO) entries scattered into a dense row-major device
// matrix: the device side of loading a sparse system from disk.  The
// reference builds a dense object array on the host instead
// (linac/sparse_matrix_tools.py:37-44, matrix_from_coo).  The caller reduces
// the values modulo p; entries not listed are zero.
#include <stddef.h>
#include <stdint.h>

#include <algorithm>

#include "gf_common.hpp"

namespace linac {

template <class T>
__global__ __launch_bounds__(256) void k_fill_zero(T* __restrict__ out, int64_t ld, int64_t rows, int64_t cols) {
    const int64_t total = rows * cols;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = x / cols;
        out[r * ld + (x - r * cols)] = (T)0;
    }
}

// out[ri[i], ci[i]] = v[i]; an index outside the matrix sets *bad instead.
template <class T>
__global__ __launch_bounds__(256) void k_coo_scatter(T* __restrict__ out, int64_t ld, int64_t rows, int64_t cols,
                                                     const int64_t* __restrict__ ri, const int64_t* __restrict__ ci,
                                                     const T* __restrict__ v, int64_t nnz, int* bad) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = ri[i], c = ci[i];
        if (r < 0 || r >= rows || c < 0 || c >= cols) {
            atomicOr(bad, 1);
            continue;
        }
        out[r * ld + c] = v[i];
    }
}

__global__ void k_clear_flag(int* p) { *p = 0; }

template <class T>
static int coo_entry(T* out, int64_t ld, int64_t rows, int64_t cols, const int64_t* ri, const int64_t* ci,
                     const T* vals, int64_t nnz, uint32_t flags) {
    if (rows < 0 || cols < 0 || nnz < 0) return fail(LINAC_ERR_INVALID, "negative size");
    if (ld < cols) return fail(LINAC_ERR_INVALID, "ld (%lld) < cols (%lld)", (long long)ld, (long long)cols);
    if (rows == 0 || cols == 0) return nnz == 0 ? LINAC_OK : fail(LINAC_ERR_INVALID, "entries for an empty matrix");
    if (!out || (nnz > 0 && (!ri || !ci || !vals))) return fail(LINAC_ERR_INVALID, "NULL operand");
    const bool on_device = (flags & LINAC_DEVICE_PTR) != 0;  // index and value arrays in device memory
    if (!on_device)
        for (int64_t i = 0; i < nnz; ++i)
            if (ri[i] < 0 || ri[i] >= rows || ci[i] < 0 || ci[i] >= cols)
                return fail(LINAC_ERR_INVALID, "entry %lld at (%lld, %lld) is outside the %lld x %lld matrix",
                            (long long)i, (long long)ri[i], (long long)ci[i], (long long)rows, (long long)cols);
    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    const size_t idx_b = (size_t)round_up(std::max<int64_t>(1, nnz) * 8, 256);
    const size_t val_b = (size_t)round_up(std::max<int64_t>(1, nnz) * (int64_t)sizeof(T), 256);
    void* p = nullptr;
    LINAC_TRY(ctx->get(SLOT_AUX, 256 + (on_device ? 0 : 2 * idx_b + val_b), &p));
    int* d_bad = (int*)p;
    const int64_t* d_ri = ri;
    const int64_t* d_ci = ci;
    const T* d_v = vals;
    if (!on_device && nnz > 0) {
        char* q = (char*)p + 256;
        d_ri = (const int64_t*)q;
        d_ci = (const int64_t*)(q + idx_b);
        d_v = (const T*)(q + 2 * idx_b);
        LINAC_HIP(hipMemcpyAsync((void*)d_ri, ri, nnz * 8, hipMemcpyHostToDevice, s));
        LINAC_HIP(hipMemcpyAsync((void*)d_ci, ci, nnz * 8, hipMemcpyHostToDevice, s));
        LINAC_HIP(hipMemcpyAsync((void*)d_v, vals, nnz * sizeof(T), hipMemcpyHostToDevice, s));
    }
    k_clear_flag<<<1, 1, 0, s>>>(d_bad);
    const unsigned gz = (unsigned)std::max<int64_t>(1, std::min<int64_t>(16384, ceil_div(rows * cols, 256)));
    k_fill_zero<T><<<gz, 256, 0, s>>>(out, ld, rows, cols);
    if (nnz > 0) {
        const unsigned gs = (unsigned)std::max<int64_t>(1, std::min<int64_t>(16384, ceil_div(nnz, 256)));
        k_coo_scatter<T><<<gs, 256, 0, s>>>(out, ld, rows, cols, d_ri, d_ci, d_v, nnz, d_bad);
    }
    LINAC_HIP(hipGetLastError());
    int bad = 0;
    LINAC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipStreamSynchronize(s));
    if (bad) return fail(LINAC_ERR_INVALID, "a COO index lies outside the %lld x %lld matrix", (long long)rows,
                         (long long)cols);
    return LINAC_OK;
}

}  // namespace linac

using namespace linac;

extern "C" {

int linac_coo_scatter_u32(uint32_t* out, int64_t ld, int64_t rows, int64_t cols, const int64_t* row_idx,
                          const int64_t* col_idx, const uint32_t* values, int64_t nnz, uint32_t flags) {
    return coo_entry(out, ld, rows, cols, row_idx, col_idx, values, nnz, flags);
}

int linac_coo_scatter_u64(uint64_t* out, int64_t ld, int64_t rows, int64_t cols, const int64_t* row_idx,
                          const int64_t* col_idx, const uint64_t* values, int64_t nnz, uint32_t flags) {
    return coo_entry(out, ld, rows, cols, row_idx, col_idx, values, nnz, flags);
}

}  // extern "C"
