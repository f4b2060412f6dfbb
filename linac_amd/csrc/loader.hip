// Monomial-evaluation matrix loader.
//
//   out[r, c] = prod_{t < degree} bases[r, lindices[c, t]]      (mod p, or in
//   double / complex double arithmetic when p == 0)
//
// Replaces cuda_load_matrix (linac/pycuda_matrix_loader.py:28-124) and its
// LoadMatrix kernel (linac/matrix_loader.cu:88-112).  Differences: the prime
// is a run-time argument (no JIT per shape), each output entry is written once
// (the reference does `degree` read-modify-writes of global memory), indices
// are validated, and a workgroup keeps the index lists of its 256 columns in
// registers while it walks ROWS_PER_BLOCK rows whose bases sit in LDS.  Many
// systems that share `lindices` (one per evaluation point of a sweep) are one
// call: stack their bases row-wise.
#include <algorithm>
#include <type_traits>

#include "field.hpp"
#include "runtime.hpp"

namespace linac {

constexpr int LOADER_COLS = 256;      // columns per workgroup (one per thread)
constexpr int LOADER_ROWS = 16;       // rows per workgroup (fewer when the bases are long)
constexpr int LOADER_LDS = 65536;     // LDS budget for the staged bases
constexpr int LOADER_MAX_DEGREE = 32; // index registers per thread
constexpr int LOADER_MAX_BASIS = 4096;

struct CplxD {
    double re, im;
};

template <class F>
struct ModArith {
    using T = typename F::T;
    F f;
    __device__ T load(T x) const { return f.to_mont(f.reduce(x)); }        // into Montgomery form
    __device__ T one() const { return f.to_mont((T)1); }
    __device__ T mul(T a, T b) const { return f.mont(a, b); }
    __device__ T store(T x) const { return f.mont(x, (T)1); }              // out of Montgomery form
};

struct F64Arith {
    using T = double;
    __device__ double load(double x) const { return x; }
    __device__ double one() const { return 1.0; }
    __device__ double mul(double a, double b) const { return a * b; }
    __device__ double store(double x) const { return x; }
};

struct C128Arith {
    using T = CplxD;
    __device__ CplxD load(CplxD x) const { return x; }
    __device__ CplxD one() const { return CplxD{1.0, 0.0}; }
    __device__ CplxD mul(CplxD a, CplxD b) const {
        return CplxD{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    }
    __device__ CplxD store(CplxD x) const { return x; }
};

template <class Ar>
__global__ __launch_bounds__(LOADER_COLS) void k_load_matrix(typename Ar::T* __restrict__ out, int64_t ldo,
                                                             const typename Ar::T* __restrict__ bases,
                                                             const uint32_t* __restrict__ lindices, int64_t rows,
                                                             int64_t cols, int basis_len, int degree, int rpb,
                                                             Ar ar, int* __restrict__ bad_index) {
    using T = typename Ar::T;
    extern __shared__ unsigned char dyn[];
    T* sb = reinterpret_cast<T*>(dyn);  // [rpb][basis_len]
    const int64_t c = (int64_t)blockIdx.x * LOADER_COLS + threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * rpb;
    const int nrows = (int)std::min<int64_t>(rpb, rows - r0);
    for (int x = threadIdx.x; x < nrows * basis_len; x += LOADER_COLS) {
        const int i = x / basis_len, j = x - i * basis_len;
        sb[i * basis_len + j] = ar.load(bases[(r0 + i) * basis_len + j]);
    }
    uint32_t idx[LOADER_MAX_DEGREE];
    bool ok = c < cols;
    if (ok) {
#pragma unroll
        for (int t = 0; t < LOADER_MAX_DEGREE; ++t) {
            if (t < degree) {
                const uint32_t v = lindices[c * degree + t];
                if (v >= (uint32_t)basis_len) {
                    atomicExch(bad_index, 1);
                    ok = false;
                }
                idx[t] = v < (uint32_t)basis_len ? v : 0u;
            }
        }
    }
    __syncthreads();
    if (c >= cols) return;
    for (int i = 0; i < nrows; ++i) {
        const T* b = sb + i * basis_len;
        T v = ar.one();
#pragma unroll
        for (int t = 0; t < LOADER_MAX_DEGREE; ++t)
            if (t < degree) v = ar.mul(v, b[idx[t]]);
        out[(r0 + i) * ldo + c] = ok ? ar.store(v) : T{};
    }
}

__global__ void k_zero_int(int* p) { *p = 0; }

template <class Ar>
static int load_entry(typename Ar::T* out, const typename Ar::T* bases, const uint32_t* lindices, int64_t rows,
                      int64_t cols, int64_t basis_len, int64_t degree, const Ar& ar, uint32_t flags,
                      void* stream_arg) {
    using T = typename Ar::T;
    if (rows < 0 || cols < 0 || basis_len < 0 || degree < 0)
        return fail(LINAC_ERR_INVALID, "negative size");
    if (rows == 0 || cols == 0) return LINAC_OK;
    if (degree == 0) return fail(LINAC_ERR_INVALID, "degree must be >= 1");
    if (degree > LOADER_MAX_DEGREE)
        return fail(LINAC_ERR_INVALID, "degree %lld exceeds %d", (long long)degree, LOADER_MAX_DEGREE);
    if (basis_len < 1 || basis_len > LOADER_MAX_BASIS)
        return fail(LINAC_ERR_INVALID, "basis length %lld outside [1, %d]", (long long)basis_len, LOADER_MAX_BASIS);
    if (!out || !bases || !lindices) return fail(LINAC_ERR_INVALID, "NULL operand");
    const bool on_device = (flags & LINAC_DEVICE_PTR) != 0;
    if (!on_device) {
        for (int64_t x = 0; x < cols * degree; ++x)
            if (lindices[x] >= (uint64_t)basis_len)
                return fail(LINAC_ERR_INVALID, "lindices[%lld] = %u is not < basis length %lld", (long long)x,
                            lindices[x], (long long)basis_len);
    }
    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipStream_t s = stream_arg ? (hipStream_t)stream_arg : ctx->stream;
    T* d_out = out;
    const T* d_bases = bases;
    const uint32_t* d_lind = lindices;
    int* d_bad = nullptr;
    const size_t out_b = (size_t)rows * cols * sizeof(T), bases_b = (size_t)rows * basis_len * sizeof(T),
                 lind_b = (size_t)cols * degree * 4;
    void* p = nullptr;
    size_t need = 256 + (on_device ? 0 : round_up((int64_t)lind_b, 256) + round_up((int64_t)bases_b, 256));
    LINAC_TRY(ctx->get(SLOT_AUX, need, &p));
    d_bad = (int*)p;
    if (!on_device) {
        char* q = (char*)p + 256;
        d_lind = (const uint32_t*)q;
        d_bases = (const T*)(q + round_up((int64_t)lind_b, 256));
        void* o = nullptr;
        LINAC_TRY(ctx->get(SLOT_MATRIX, out_b, &o));
        d_out = (T*)o;
        LINAC_HIP(hipMemcpyAsync((void*)d_lind, lindices, lind_b, hipMemcpyHostToDevice, s));
        LINAC_HIP(hipMemcpyAsync((void*)d_bases, bases, bases_b, hipMemcpyHostToDevice, s));
    }
    k_zero_int<<<1, 1, 0, s>>>(d_bad);
    const int rpb = (int)std::max<int64_t>(1, std::min<int64_t>(LOADER_ROWS, LOADER_LDS / (basis_len * (int64_t)sizeof(T))));
    const size_t lds = (size_t)rpb * basis_len * sizeof(T);
    const int64_t max_rows_per_launch = (int64_t)65535 * rpb;
    for (int64_t r0 = 0; r0 < rows; r0 += max_rows_per_launch) {
        const int64_t nr = std::min<int64_t>(rows - r0, max_rows_per_launch);
        dim3 grid((unsigned)ceil_div(cols, LOADER_COLS), (unsigned)ceil_div(nr, rpb));
        k_load_matrix<Ar><<<grid, LOADER_COLS, lds, s>>>(d_out + r0 * cols, cols, d_bases + r0 * basis_len, d_lind, nr,
                                                          cols, (int)basis_len, (int)degree, rpb, ar, d_bad);
    }
    LINAC_HIP(hipGetLastError());
    int bad = 0;
    LINAC_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    if (!on_device) LINAC_HIP(hipMemcpyAsync(out, d_out, out_b, hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipStreamSynchronize(s));
    if (bad) return fail(LINAC_ERR_INVALID, "an index in lindices is not below the basis length");
    return LINAC_OK;
}

}  // namespace linac

using namespace linac;

extern "C" {

int linac_load_matrix_u32(uint32_t* out, const uint32_t* bases, const uint32_t* lindices, int64_t rows, int64_t cols,
                          int64_t basis_len, int64_t degree, uint32_t p, uint32_t flags, void* stream) {
    if (!is_prime_u64(p)) return fail(LINAC_ERR_INVALID, "field characteristic %u is not prime", p);
    if (p == 2)
        return load_entry(out, bases, lindices, rows, cols, basis_len, degree, ModArith<FieldGF2>{FieldGF2{2u}},
                          flags, stream);
    return load_entry(out, bases, lindices, rows, cols, basis_len, degree, ModArith<FieldU32>{make_field_u32(p)},
                      flags, stream);
}

int linac_load_matrix_u64(uint64_t* out, const uint64_t* bases, const uint32_t* lindices, int64_t rows, int64_t cols,
                          int64_t basis_len, int64_t degree, uint64_t p, uint32_t flags, void* stream) {
    if (p >= (1ull << 63)) return fail(LINAC_ERR_INVALID, "field characteristic >= 2^63 is not supported");
    if (!is_prime_u64(p) || p == 2)
        return fail(LINAC_ERR_INVALID, "field characteristic %llu is not an odd prime", (unsigned long long)p);
    return load_entry(out, bases, lindices, rows, cols, basis_len, degree, ModArith<FieldU64>{make_field_u64(p)},
                      flags, stream);
}

int linac_load_matrix_f64(double* out, const double* bases, const uint32_t* lindices, int64_t rows, int64_t cols,
                          int64_t basis_len, int64_t degree, uint32_t flags, void* stream) {
    return load_entry(out, bases, lindices, rows, cols, basis_len, degree, F64Arith{}, flags, stream);
}

int linac_load_matrix_c128(double* out, const double* bases, const uint32_t* lindices, int64_t rows, int64_t cols,
                           int64_t basis_len, int64_t degree, uint32_t flags, void* stream) {
    return load_entry((CplxD*)out, (const CplxD*)bases, lindices, rows, cols, basis_len, degree, C128Arith{}, flags,
                      stream);
}

}  // extern "C"
