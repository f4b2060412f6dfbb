// Residual-based refinement of the float RREF's solution columns (see
// float_refine.hpp for the method).  Kernels:
//   k_gather_x         Xc = X (the solution in the pivot rows), compact;
//   k_refine_residual  E = (Aorig[:, NP] - Aorig[:, P] X) / scale, summed in
//   (_narrow for few   double-double (Dekker's exact product by FMA, Knuth's
//   solution columns)  two-sum), rounded to float64 once per entry;
//   k_replay_panel     one panel: Y = ZQ_i E[Q_i, :], E[Q_i[t]] = Y[t],
//                      E[r] += ZL_i[r] Y (ping-pong buffers);
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
    DD v;  // no initialiser: the type also lives in LDS (init() sets it)
    __device__ void init(double c) { v = DD{c, 0.0}; }
    __device__ void sub_prod(double a, double x) { v = dd_add_prod(v, -a, x); }
    __device__ void merge(const DDAcc& o) { v = dd_add(v, o.v); }
    __device__ void add_value(double c) { v = dd_add(v, DD{c, 0.0}); }
    __device__ double result(double scale) const { return (v.hi + v.lo) / scale; }
};
template <>
struct DDAcc<CplxF> {
    DD re, im;
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
    __device__ void merge(const DDAcc& o) {
        re = dd_add(re, o.re);
        im = dd_add(im, o.im);
    }
    __device__ void add_value(CplxF c) {
        re = dd_add(re, DD{c.re, 0.0});
        im = dd_add(im, DD{c.im, 0.0});
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

// Xc[t, j] = A[PR[t], NP[j]]: the current solution, compact (rank x nr).
template <class T>
__global__ __launch_bounds__(256) void k_gather_x(const T* __restrict__ A, int64_t ld, const int64_t* __restrict__ PR,
                                                  int64_t rank, const int64_t* __restrict__ NP, int64_t nr,
                                                  T* __restrict__ Xc) {
    const int64_t total = rank * nr;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = x / nr, j = x - t * nr;
        Xc[x] = A[PR[t] * ld + NP[j]];
    }
}

// 64 rows x 64 solution columns per workgroup; pivots in chunks of 16 through
// LDS; each thread owns 4 x 4 outputs.  For many solution columns.
constexpr int RR_M = 64, RR_N = 64, RR_K = 16;

template <class T>
__global__ __launch_bounds__(256) void k_refine_residual(const T* __restrict__ Ao, int64_t ldo, const T* __restrict__ Xc,
                                                         const int64_t* __restrict__ PC, int64_t rank,
                                                         const int64_t* __restrict__ NP, int64_t nr, int64_t N,
                                                         const double* __restrict__ scale, T* __restrict__ E) {
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
            sX[k][jj] = (t < rank && j < nr) ? Xc[t * nr + j] : FloatOps<T>::zero();
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
            if (r < N && j < nr) E[r * nr + j] = acc[a][b].result(scale[r]);
        }
}

// The same for at most RN_MAX solution columns: one workgroup per row, the
// pivots split over its threads, partial double-double sums combined in LDS.
constexpr int RN_MAX = 4;

template <class T>
__global__ __launch_bounds__(256) void k_refine_residual_narrow(const T* __restrict__ Ao, int64_t ldo,
                                                                const T* __restrict__ Xc,
                                                                const int64_t* __restrict__ PC, int64_t rank,
                                                                const int64_t* __restrict__ NP, int nr,
                                                                const double* __restrict__ scale, T* __restrict__ E) {
    __shared__ DDAcc<T> part[256];
    const int64_t r = blockIdx.x;
    const T* row = Ao + r * ldo;
    for (int j0 = 0; j0 < nr; j0 += RN_MAX) {
        DDAcc<T> acc[RN_MAX];
#pragma unroll
        for (int q = 0; q < RN_MAX; ++q) acc[q].init(FloatOps<T>::zero());
        for (int64_t t = threadIdx.x; t < rank; t += 256) {
            const T a = row[PC[t]];
#pragma unroll
            for (int q = 0; q < RN_MAX; ++q)
                if (j0 + q < nr) acc[q].sub_prod(a, Xc[t * nr + j0 + q]);
        }
#pragma unroll
        for (int q = 0; q < RN_MAX; ++q) {
            if (j0 + q >= nr) break;
            __syncthreads();
            part[threadIdx.x] = acc[q];
            __syncthreads();
            for (int w = 128; w > 0; w >>= 1) {
                if (threadIdx.x < w) part[threadIdx.x].merge(part[threadIdx.x + w]);
                __syncthreads();
            }
            if (threadIdx.x == 0) {
                DDAcc<T> tot = part[0];
                tot.add_value(row[NP[j0 + q]]);
                E[r * nr + j0 + q] = tot.result(scale[r]);
            }
        }
    }
}

// One panel of the replay, Ein -> Eout (so that every workgroup reads the
// panel's pivot rows before any is overwritten):
//   Y = ZQ Ein[Q, :],  Eout[Q[t]] = Y[t],  Eout[r] = Ein[r] + ZL[r] Y  (other rows).
// A workgroup takes RP_ROWS rows and JW columns (JW = nr rounded up to a
// power of two, at most 64) and forms its columns of Y itself.
constexpr int RP_ROWS = 64;

template <class T>
__global__ __launch_bounds__(256) void k_replay_panel(const T* __restrict__ ZL, const T* __restrict__ ZQ,
                                                      const int64_t* __restrict__ Q, int k,
                                                      const int32_t* __restrict__ pepoch, int panel,
                                                      const T* __restrict__ Ein, T* __restrict__ Eout, int64_t nr,
                                                      int64_t N, int jw) {
    __shared__ T sE[RF_FB][64 + 1];
    __shared__ T sY[RF_FB][64 + 1];
    __shared__ int64_t sQ[RF_FB];
    const int64_t j0 = (int64_t)blockIdx.y * jw;
    if (threadIdx.x < k) sQ[threadIdx.x] = Q[threadIdx.x];
    __syncthreads();
    for (int x = threadIdx.x; x < k * jw; x += 256) {
        const int sidx = x / jw, jj = x % jw;
        sE[sidx][jj] = j0 + jj < nr ? Ein[sQ[sidx] * nr + j0 + jj] : FloatOps<T>::zero();
    }
    __syncthreads();
    for (int x = threadIdx.x; x < k * jw; x += 256) {
        const int t = x / jw, jj = x % jw;
        T y = FloatOps<T>::zero();
        const T* zq = ZQ + t * RF_FB;
        for (int sidx = 0; sidx < k; ++sidx) y = FloatOps<T>::fms(y, FloatOps<T>::neg(zq[sidx]), sE[sidx][jj]);
        sY[t][jj] = y;
    }
    __syncthreads();
    const int rows_per_pass = 256 / jw;
    const int jj = threadIdx.x % jw;
    const int64_t j = j0 + jj;
    for (int64_t r = (int64_t)blockIdx.x * RP_ROWS + threadIdx.x / jw; r < min(N, (int64_t)(blockIdx.x + 1) * RP_ROWS);
         r += rows_per_pass) {
        if (j >= nr) continue;
        if (pepoch[r] == panel) {
            int t = 0;
            while (t < k && sQ[t] != r) ++t;
            Eout[r * nr + j] = t < k ? sY[t][jj] : Ein[r * nr + j];
            continue;
        }
        T e = Ein[r * nr + j];
        const T* zl = ZL + r * RF_FB;
        for (int t = 0; t < k; ++t) e = FloatOps<T>::fms(e, FloatOps<T>::neg(zl[t]), sY[t][jj]);
        Eout[r * nr + j] = e;
    }
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
                 hipStream_t s, DeviceContext* ctx, double stop_below) {
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
    // scratch: NP list, row scales, stats, Xc (rank x nr), E twice (N x nr)
    auto al = [](size_t x) { return (size_t)round_up((int64_t)x, 256); };
    const size_t o_np = 0, o_sc = o_np + al(nr * 8), o_st = o_sc + al(N * 8), o_x = o_st + 256,
                 o_e0 = o_x + al((size_t)rank * nr * sizeof(T)), o_e1 = o_e0 + al((size_t)N * nr * sizeof(T)),
                 total = o_e1 + al((size_t)N * nr * sizeof(T));
    void* base = nullptr;
    LINAC_TRY(ctx->get(SLOT_OUT, total, &base));
    char* b = (char*)base;
    int64_t* d_np = (int64_t*)(b + o_np);
    double* d_scale = (double*)(b + o_sc);
    unsigned long long* d_stats = (unsigned long long*)(b + o_st);
    T* Xc = (T*)(b + o_x);
    T* Ebuf[2] = {(T*)(b + o_e0), (T*)(b + o_e1)};
    LINAC_HIP(hipMemcpyAsync(d_np, np.data(), nr * 8, hipMemcpyHostToDevice, s));
    k_refine_scales<T><<<(unsigned)N, 256, 0, s>>>(Aorig, ldo, M, d_scale);
    int jw = 1;  // columns per replay workgroup: nr rounded up to a power of two, at most 64
    while (jw < nr && jw < 64) jw *= 2;
    const unsigned gx = (unsigned)std::min<int64_t>(4096, ceil_div(rank * nr, 256));
    for (int pass = 0; pass < passes; ++pass) {
        k_gather_x<T><<<gx, 256, 0, s>>>(A, ld, piv_rows, rank, d_np, nr, Xc);
        if (nr <= 4 * RN_MAX)
            k_refine_residual_narrow<T><<<(unsigned)N, 256, 0, s>>>(Aorig, ldo, Xc, piv_cols, rank, d_np, (int)nr,
                                                                     d_scale, Ebuf[0]);
        else
            k_refine_residual<T><<<dim3((unsigned)ceil_div(nr, RR_N), (unsigned)ceil_div(N, RR_M)), 256, 0, s>>>(
                Aorig, ldo, Xc, piv_cols, rank, d_np, nr, N, d_scale, Ebuf[0]);
        int cur = 0;
        for (int64_t i = 0; i < fac.npanels; ++i) {
            const int k = fac.kp[(size_t)i];
            if (k == 0) continue;
            const T* ZQ = (const T*)fac.ZQ + i * RF_FB * RF_FB;
            const T* ZL = (const T*)fac.ZL + i * N * RF_FB;
            const int64_t* Q = fac.qrows + i * RF_FB;
            k_replay_panel<T><<<dim3((unsigned)ceil_div(N, RP_ROWS), (unsigned)ceil_div(nr, jw)), 256, 0, s>>>(
                ZL, ZQ, Q, k, fac.pepoch, (int)i, Ebuf[cur], Ebuf[cur ^ 1], nr, N, jw);
            cur ^= 1;
        }
        LINAC_HIP(hipMemsetAsync(d_stats, 0, 16, s));
        k_refine_apply<T><<<gx, 256, 0, s>>>(A, ld, piv_rows, rank, d_np, nr, Ebuf[cur], nr, d_stats);
        LINAC_HIP(hipGetLastError());
        unsigned long long st[2] = {0, 0};
        LINAC_HIP(hipMemcpyAsync(st, d_stats, 16, hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
        double md, mx;
        __builtin_memcpy(&md, &st[0], 8);
        __builtin_memcpy(&mx, &st[1], 8);
        const double ratio = mx > 0.0 ? md / mx : 0.0;
        const double prev = rs.last_correction;
        if (pass == 0) rs.first_correction = ratio;
        rs.last_correction = ratio;
        rs.iterations = pass + 1;
        if (stop_below > 0.0 && (ratio <= stop_below || (pass > 0 && ratio > 0.5 * prev))) break;
    }
    return LINAC_OK;
}

template int float_refine<double>(double*, int64_t, const double*, int64_t, int64_t, int64_t, const int64_t*,
                                  const int64_t*, const std::vector<int64_t>&, const FloatFactors&, int, hipStream_t,
                                  DeviceContext*, double);
template int float_refine<CplxF>(CplxF*, int64_t, const CplxF*, int64_t, int64_t, int64_t, const int64_t*,
                                 const int64_t*, const std::vector<int64_t>&, const FloatFactors&, int, hipStream_t,
                                 DeviceContext*, double);

}  // namespace linac

using namespace linac;

extern "C" {

int linac_refine_last(double* out, int n) {
    if (!out || n <= 0) return 0;
    const RefineStats& rs = thread_refine_stats();
    const double v[5] = {rs.iterations, rs.first_correction, rs.last_correction, rs.solution_columns, rs.fallback};
    const int m = n < 5 ? n : 5;
    for (int i = 0; i < m; ++i) out[i] = v[i];
    return m;
}

}  // extern "C"
