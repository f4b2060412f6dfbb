// Residual-based refinement of the float RREF's solution columns (see
// float_refine.hpp for the method).  Kernels:
//   k_refine_residual  E = (Aorig[:, NP] - Aorig[:, P] X) / scale, summed in
//                      double-double (Dekker's exact product by FMA, Knuth's
//                      two-sum), rounded to float64 once per entry;
//   k_replay_y         Y = ZQ_i E[Q_i, :]          (one panel's pivot rows)
//   k_replay_rows      E[r] += ZL_i[r] Y, E[Q_i[t]] = Y[t]
//   k_refine_apply     X += E[Q, :], and max |E[Q, :]|, max |X| for the stats.
#include <algorithm>
#include <vector>

#include "float_ops.hpp"
#include "float_refine.hpp"

namespace linac {

constexpr int RF_FB = 64;  // panel width of the blocked float sweep

RefineStats& thread_refine_stats() {
    static thread_local RefineStats st;
    return st;
}

// ---- double-double accumulation (no contraction: every rounding is the
// one the error-free transformations account for) ----
struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_add_prod(DD acc, double a, double b) {
#pragma clang fp contract(off)
    const double p = a * b;
    const double pe = fma(a, b, -p);  // exact: a b = p + pe
    const double s = acc.hi + p;
    const double bb = s - acc.hi;
    const double se = (acc.hi - (s - bb)) + (p - bb);  // exact: acc.hi + p = s + se
    const double t = se + (acc.lo + pe);
    const double hi = s + t;
    return DD{hi, t - (hi - s)};
}

__device__ __forceinline__ DD dd_add(DD x, DD y) {
#pragma clang fp contract(off)
    const double s = x.hi + y.hi;
    const double bb = s - x.hi;
    const double se = (x.hi - (s - bb)) + (y.hi - bb);
    const double t = se + (x.lo + y.lo);
    const double hi = s + t;
    return DD{hi, t - (hi - s)};
}

// Element-type glue: a T accumulator of double-doubles.
template <class T>
struct DDAcc;
template <>
struct DDAcc<double> {
    DD v{0.0, 0.0};
    __device__ void init(double c) { v = DD{c, 0.0}; }
    __device__ void sub_prod(double a, double x) { v = dd_add_prod(v, -a, x); }
    __device__ double result(double scale) const { return (v.hi + v.lo) / scale; }
};
template <>
struct DDAcc<CplxF> {
    DD re{0.0, 0.0}, im{0.0, 0.0};
    __device__ void init(CplxF c) {
        re = DD{c.re, 0.0};
        im = DD{c.im, 0.0};
    }
    __device__ void sub_prod(CplxF a, CplxF x) {  // -= a x
        re = dd_add_prod(re, -a.re, x.re);
        re = dd_add_prod(re, a.im, x.im);
        im = dd_add_prod(im, -a.re, x.im);
        im = dd_add_prod(im, -a.im, x.re);
    }
    __device__ CplxF result(double scale) const { return CplxF{(re.hi + re.lo) / scale, (im.hi + im.lo) / scale}; }
};

__device__ __forceinline__ double mag(double x) { return fabs(x); }
__device__ __forceinline__ double mag(CplxF z) { return fmax(fabs(z.re), fabs(z.im)); }

// Row scales of the sweep: max |Aorig[r, :]| (1 for an all-zero row, which
// the sweep leaves unscaled).
template <class T>
__global__ __launch_bounds__(256) void k_refine_scales(const T* __restrict__ Ao, int64_t ldo, int64_t M,
                                                       double* __restrict__ scale) {
    __shared__ double s_max[256];
    const int64_t r = blockIdx.x;
    double m = 0.0;
    for (int64_t c = threadIdx.x; c < M; c += 256) m = fmax(m, FloatOps<T>::abs(Ao[r * ldo + c]));
    s_max[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[r] = s_max[0] == 0.0 ? 1.0 : s_max[0];
}

// 64 rows x 64 solution columns per workgroup; pivots in chunks of 16 through
// LDS; each thread owns 4 x 4 outputs.
constexpr int RR_M = 64, RR_N = 64, RR_K = 16;

template <class T>
__global__ __launch_bounds__(256) void k_refine_residual(const T* __restrict__ Ao, int64_t ldo, const T* __restrict__ A,
                                                         int64_t ld, const int64_t* __restrict__ PR,
                                                         const int64_t* __restrict__ PC, int64_t rank,
                                                         const int64_t* __restrict__ NP, int64_t nr, int64_t N,
                                                         const double* __restrict__ scale, T* __restrict__ E,
                                                         int64_t lde) {
    __shared__ T sA[RR_M][RR_K + 1];
    __shared__ T sX[RR_K][RR_N + 1];
    const int64_t r0 = (int64_t)blockIdx.y * RR_M, j0 = (int64_t)blockIdx.x * RR_N;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;  // outputs (ty + 16 a, tx + 16 b)
    DDAcc<T> acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t r = r0 + ty + 16 * a, j = j0 + tx + 16 * b;
            acc[a][b].init(r < N && j < nr ? Ao[r * ldo + NP[j]] : FloatOps<T>::zero());
        }
    for (int64_t t0 = 0; t0 < rank; t0 += RR_K) {
        for (int x = threadIdx.x; x < RR_M * RR_K; x += 256) {
            const int i = x / RR_K, k = x % RR_K;
            const int64_t r = r0 + i, t = t0 + k;
            sA[i][k] = (r < N && t < rank) ? Ao[r * ldo + PC[t]] : FloatOps<T>::zero();
        }
        for (int x = threadIdx.x; x < RR_K * RR_N; x += 256) {
            const int k = x / RR_N, jj = x % RR_N;
            const int64_t t = t0 + k, j = j0 + jj;
            sX[k][jj] = (t < rank && j < nr) ? A[PR[t] * ld + NP[j]] : FloatOps<T>::zero();
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < RR_K; ++k)
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b].sub_prod(sA[ty + 16 * a][k], sX[k][tx + 16 * b]);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int64_t r = r0 + ty + 16 * a, j = j0 + tx + 16 * b;
            if (r < N && j < nr) E[r * lde + j] = acc[a][b].result(scale[r]);
        }
}

// Y[t, j] = sum_{s<k} ZQ[t][s] E[Q[s], j] for t < k; 64 columns per workgroup.
template <class T>
__global__ __launch_bounds__(256) void k_replay_y(const T* __restrict__ ZQ, const int64_t* __restrict__ Q, int k,
                                                  const T* __restrict__ E, int64_t lde, int64_t nr, T* __restrict__ Y) {
    __shared__ T sE[RF_FB][64 + 1];
    const int64_t j0 = (int64_t)blockIdx.x * 64;
    for (int x = threadIdx.x; x < k * 64; x += 256) {
        const int sidx = x / 64, jj = x % 64;
        sE[sidx][jj] = j0 + jj < nr ? E[Q[sidx] * lde + j0 + jj] : FloatOps<T>::zero();
    }
    __syncthreads();
    for (int x = threadIdx.x; x < k * 64; x += 256) {
        const int t = x / 64, jj = x % 64;
        if (j0 + jj >= nr) continue;
        T y = FloatOps<T>::zero();
        for (int sidx = 0; sidx < k; ++sidx) y = FloatOps<T>::fms(y, FloatOps<T>::neg(ZQ[t * RF_FB + sidx]), sE[sidx][jj]);
        Y[(int64_t)t * nr + j0 + jj] = y;
    }
}

// E[r, j] += sum_{t<k} ZL[r][t] Y[t, j] for rows that are not pivots of this
// panel (their ZL rows are zero), E[Q[t], j] = Y[t, j] for those that are.
// Four rows x 64 columns per workgroup.
template <class T>
__global__ __launch_bounds__(256) void k_replay_rows(const T* __restrict__ ZL, const int64_t* __restrict__ Q, int k,
                                                     const int32_t* __restrict__ pepoch, int panel, const T* __restrict__ Y,
                                                     T* __restrict__ E, int64_t lde, int64_t nr, int64_t N) {
    __shared__ T sY[RF_FB][64 + 1];
    __shared__ int64_t sQ[RF_FB];
    const int64_t j0 = (int64_t)blockIdx.y * 64;
    for (int x = threadIdx.x; x < k * 64; x += 256) {
        const int t = x / 64, jj = x % 64;
        sY[t][jj] = j0 + jj < nr ? Y[(int64_t)t * nr + j0 + jj] : FloatOps<T>::zero();
    }
    if (threadIdx.x < k) sQ[threadIdx.x] = Q[threadIdx.x];
    __syncthreads();
    const int jj = threadIdx.x & 63;
    const int64_t j = j0 + jj;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= N || j >= nr) return;
    if (pepoch[r] == panel) {
        int t = 0;
        while (t < k && sQ[t] != r) ++t;
        if (t < k) E[r * lde + j] = sY[t][jj];
        return;
    }
    T e = E[r * lde + j];
    const T* zl = ZL + r * RF_FB;
    for (int t = 0; t < k; ++t) e = FloatOps<T>::fms(e, FloatOps<T>::neg(zl[t]), sY[t][jj]);
    E[r * lde + j] = e;
}

// X[t, j] += E[Q[t], j] (X in place in A's pivot rows); stats[0] = max |E[Q]|,
// stats[1] = max |X| as non-negative double bit patterns.
template <class T>
__global__ __launch_bounds__(256) void k_refine_apply(T* __restrict__ A, int64_t ld, const int64_t* __restrict__ PR,
                                                      int64_t rank, const int64_t* __restrict__ NP, int64_t nr,
                                                      const T* __restrict__ E, int64_t lde,
                                                      unsigned long long* __restrict__ stats) {
    double md = 0.0, mx = 0.0;
    const int64_t total = rank * nr;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = x / nr, j = x - t * nr;
        const T d = E[PR[t] * lde + j];
        T& v = A[PR[t] * ld + NP[j]];
        v = FloatOps<T>::fms(v, FloatOps<T>::neg(d), FloatOps<T>::one());
        md = fmax(md, mag(d));
        mx = fmax(mx, mag(v));
    }
    for (int off = 32; off > 0; off >>= 1) {
        md = fmax(md, __shfl_xor(md, off));
        mx = fmax(mx, __shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&stats[0], (unsigned long long)__double_as_longlong(md));
        atomicMax(&stats[1], (unsigned long long)__double_as_longlong(mx));
    }
}

template <class T>
int float_refine(T* A, int64_t ld, const T* Aorig, int64_t ldo, int64_t N, int64_t M, const int64_t* piv_rows,
                 const int64_t* piv_cols, const std::vector<int64_t>& pc_host, const FloatFactors& fac, int passes,
                 hipStream_t s, DeviceContext* ctx) {
    RefineStats& rs = thread_refine_stats();
    rs = RefineStats{};
    const int64_t rank = (int64_t)pc_host.size();
    std::vector<int64_t> np;
    {
        std::vector<char> isp((size_t)M, 0);
        for (int64_t c : pc_host) isp[(size_t)c] = 1;
        for (int64_t c = 0; c < M; ++c)
            if (!isp[(size_t)c]) np.push_back(c);
    }
    const int64_t nr = (int64_t)np.size();
    rs.solution_columns = (double)nr;
    if (nr == 0 || rank == 0 || passes <= 0) return LINAC_OK;
    // scratch: NP list, row scales, stats, E (N x nr), Y (FB x nr)
    auto al = [](size_t x) { return (size_t)round_up((int64_t)x, 256); };
    const size_t o_np = 0, o_sc = o_np + al(nr * 8), o_st = o_sc + al(N * 8), o_e = o_st + 256,
                 o_y = o_e + al((size_t)N * nr * sizeof(T)), total = o_y + al((size_t)RF_FB * nr * sizeof(T));
    void* base = nullptr;
    LINAC_TRY(ctx->get(SLOT_OUT, total, &base));
    char* b = (char*)base;
    int64_t* d_np = (int64_t*)(b + o_np);
    double* d_scale = (double*)(b + o_sc);
    unsigned long long* d_stats = (unsigned long long*)(b + o_st);
    T* E = (T*)(b + o_e);
    T* Y = (T*)(b + o_y);
    LINAC_HIP(hipMemcpyAsync(d_np, np.data(), nr * 8, hipMemcpyHostToDevice, s));
    k_refine_scales<T><<<(unsigned)N, 256, 0, s>>>(Aorig, ldo, M, d_scale);
    const dim3 rgrid((unsigned)ceil_div(nr, RR_N), (unsigned)ceil_div(N, RR_M));
    for (int pass = 0; pass < passes; ++pass) {
        k_refine_residual<T><<<rgrid, 256, 0, s>>>(Aorig, ldo, A, ld, piv_rows, piv_cols, rank, d_np, nr, N, d_scale, E,
                                                    nr);
        for (int64_t i = 0; i < fac.npanels; ++i) {
            const int k = fac.kp[(size_t)i];
            if (k == 0) continue;
            const T* ZQ = (const T*)fac.ZQ + i * RF_FB * RF_FB;
            const T* ZL = (const T*)fac.ZL + i * N * RF_FB;
            const int64_t* Q = fac.qrows + i * RF_FB;
            k_replay_y<T><<<(unsigned)ceil_div(nr, 64), 256, 0, s>>>(ZQ, Q, k, E, nr, nr, Y);
            k_replay_rows<T><<<dim3((unsigned)ceil_div(N, 4), (unsigned)ceil_div(nr, 64)), 256, 0, s>>>(
                ZL, Q, k, fac.pepoch, (int)i, Y, E, nr, nr, N);
        }
        LINAC_HIP(hipMemsetAsync(d_stats, 0, 16, s));
        k_refine_apply<T><<<(unsigned)std::min<int64_t>(1024, ceil_div(rank * nr, 256)), 256, 0, s>>>(
            A, ld, piv_rows, rank, d_np, nr, E, nr, d_stats);
        LINAC_HIP(hipGetLastError());
        unsigned long long st[2] = {0, 0};
        LINAC_HIP(hipMemcpyAsync(st, d_stats, 16, hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
        double md, mx;
        __builtin_memcpy(&md, &st[0], 8);
        __builtin_memcpy(&mx, &st[1], 8);
        const double ratio = mx > 0.0 ? md / mx : 0.0;
        if (pass == 0) rs.first_correction = ratio;
        rs.last_correction = ratio;
        rs.iterations = pass + 1;
    }
    return LINAC_OK;
}

template int float_refine<double>(double*, int64_t, const double*, int64_t, int64_t, int64_t, const int64_t*,
                                  const int64_t*, const std::vector<int64_t>&, const FloatFactors&, int, hipStream_t,
                                  DeviceContext*);
template int float_refine<CplxF>(CplxF*, int64_t, const CplxF*, int64_t, int64_t, int64_t, const int64_t*,
                                 const int64_t*, const std::vector<int64_t>&, const FloatFactors&, int, hipStream_t,
                                 DeviceContext*);

}  // namespace linac

using namespace linac;

extern "C" {

int linac_refine_last(double* out, int n) {
    if (!out || n <= 0) return 0;
    const RefineStats& rs = thread_refine_stats();
    const double v[4] = {rs.iterations, rs.first_correction, rs.last_correction, rs.solution_columns};
    const int m = n < 4 ? n : 4;
    for (int i = 0; i < m; ++i) out[i] = v[i];
    return m;
}

}  // extern "C"
