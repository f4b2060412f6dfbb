// Blocked Gauss-Jordan RREF in float64 / complex128 for gfx950.
//
// Same definition as the unblocked float path (linac/row_reduce.cu:201-270,
// 287-342, 404-418 driven by linac/pycuda_row_reduce.py:87-131): rows scaled
// once by their max |x|; per column the unused row of largest |x| (smallest
// row on ties) is the pivot iff |x| > tol; every other row is reduced by it.
// What changes is the order of the work:
//
//  * k_fpanel factorises a panel of FB = 64 columns over ALL N rows with exact
//    partial pivoting.  It is a cooperative kernel: every workgroup keeps 64
//    rows of the panel (and of Z, below) in registers, one column per lane;
//    per column each workgroup publishes its best candidate row, one grid
//    barrier, then every workgroup picks the same global winner (max |x|,
//    smallest row on ties) and updates its rows.  Pivot rows are normalised
//    in FP64 as they are found.
//  * Every row operation is also applied to Z (N x FB), in which a row that
//    becomes the t-th pivot of the panel starts as e_t, so that Z[Q] is the
//    inverse of the panel's pivot block (Q = pivot rows, pc = their columns,
//    in slot order).  For the trailing columns T:
//      pivot rows:      Y = new A[Q, T] = Z[Q] A[Q, T]
//      all other rows:  new A[r, T] = A[r, T] - P[r, pc] Y
//    where P is the panel as it was before the panel's elimination (row r's
//    multipliers).  This is the LU-style form: its multipliers are bounded by
//    the pivoting, whereas folding the inverse into the multipliers
//    (Z[r] A[Q, T]) loses accuracy through cancellation.
//  * k_fgemm applies both products with FP64 matrix cores
//    (v_mfma_f64_16x16x4_f64, FP64 accumulation), pivot rows first.  Complex
//    matrices use the same real GEMM: the left operand viewed as interleaved
//    (re, im) doubles times the real 2k x 2nT operand [[Br, Bi], [-Bi, Br]]
//    built by k_fgather.
//
// No host synchronisation between panels: the pivot count of a panel stays on
// the device and sizes the GEMM's depth there.
#include <type_traits>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "float_blocked.hpp"
#include "float_ops.hpp"
#include "float_refine.hpp"
#include "runtime.hpp"

namespace linac {

constexpr int FB = 64;      // panel width: one column per lane
constexpr int FWAVES = 4;   // waves per panel workgroup


constexpr uint64_t STAMP = 1ull << 63;
constexpr uint64_t ROW_NONE = (1ull << 62) - 1;



__device__ __forceinline__ double bcast(double x, int src) {
    const uint64_t u = __double_as_longlong(x);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ CplxF bcast(CplxF x, int src) { return CplxF{bcast(x.re, src), bcast(x.im, src)}; }
__device__ __forceinline__ double shfl_t(double x, int src) { return __shfl(x, src); }
__device__ __forceinline__ CplxF shfl_t(CplxF x, int src) { return CplxF{__shfl(x.re, src), __shfl(x.im, src)}; }

__device__ __forceinline__ double cload(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void cstore(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double cload_t(const double* p, double) { return cload(p); }
__device__ __forceinline__ CplxF cload_t(const CplxF* p, CplxF) {
    return CplxF{cload(&p->re), cload(&p->im)};
}
__device__ __forceinline__ void cstore_t(double* p, double v) { cstore(p, v); }
__device__ __forceinline__ void cstore_t(CplxF* p, CplxF v) {
    cstore(&p->re, v.re);
    cstore(&p->im, v.im);
}

__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// (val, row) ordering of candidates: larger |x| first, then smaller row;
// row < 0 marks "no candidate".
__device__ __forceinline__ bool cand_better(double v2, int64_t r2, double v1, int64_t r1) {
    if (r2 < 0) return false;
    if (r1 < 0) return true;
    return v2 > v1 || (v2 == v1 && r2 < r1);
}

template <class T, int RW>
__global__ __launch_bounds__(256) void k_fpanel(FPanelArgs<T> P) {
    using Ops = FloatOps<T>;
    // wv through readfirstlane: the compiler then knows it is wave-uniform and
    // branches on it stay scalar
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int G = gridDim.x, g = blockIdx.x;
    const int64_t row0 = ((int64_t)g * FWAVES + wv) * RW;
    __shared__ double s_val[FWAVES];
    __shared__ int s_idx[FWAVES];
    __shared__ int s_pcol[FB];  // panel column of each pivot slot
    __shared__ int s_slot[FWAVES * RW];  // pivot slot of each of this workgroup's rows (-1: none)
    __shared__ double s_win_val;
    __shared__ int64_t s_win_row;
    __shared__ int s_win_q;

    // Rows past N hold zeros and are never candidates, so the per-row loops
    // below run without branches (divergent per-row control flow makes the
    // compiler copy the whole register arrays at every merge).
    T pr[RW], zr[RW];
    uint32_t valid = 0, live = 0, pmask = 0;
    if (lane < RW) s_slot[wv * RW + lane] = -1;
    const int lc = lane < P.w ? lane : P.w - 1;  // clamped, in-bounds column for the loads
#pragma unroll
    for (int i = 0; i < RW; ++i) {
        const int64_t r = row0 + i;
        const int64_t rc = r < P.N ? r : P.N - 1;  // clamped, in-bounds row
        const T x = P.A[rc * P.ld + P.J + lc];
        pr[i] = Ops::sel(r < P.N && lane < P.w, x, Ops::zero());
        zr[i] = Ops::zero();
        const bool in = r < P.N;
        valid |= in ? (1u << i) : 0u;
        live |= (in && !P.used[rc]) ? (1u << i) : 0u;
    }
    const int64_t rank0 = *P.rank;
    int k = 0;
    for (int c = 0; c < P.w; ++c) {
        const unsigned long long t0 = wall_clock64();
        // 1. this wave's best unused row in column c (lane c holds the column)
        double best = -1.0;
        int bi = -1;
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const double v = ((live >> i) & 1u) ? Ops::abs(pr[i]) : -1.0;
            bi = v > best ? i : bi;
            best = v > best ? v : best;
        }
        best = bcast(best, c);
        bi = __builtin_amdgcn_readlane(bi, c);
        if (lane == 0) {
            s_val[wv] = best;
            s_idx[wv] = bi;
        }
        __syncthreads();
        double wb = -1.0;
        int64_t wr = -1;
        int ww = -1;
#pragma unroll
        for (int q = 0; q < FWAVES; ++q) {
            const int64_t rq = s_idx[q] >= 0 ? ((int64_t)g * FWAVES + q) * RW + s_idx[q] : -1;
            if (cand_better(s_val[q], rq, wb, wr)) {
                wb = s_val[q];
                wr = rq;
                ww = q;
            }
        }
        const unsigned long long t1 = wall_clock64();
        // 2. publish this workgroup's candidate: the row (panel | Z) first, then,
        //    after a release fence, the stamped slot.  Buffers alternate by
        //    column; the stamp is bit 1 of the global column index.
        const int64_t gcol = P.J + c;
        const int pb = (int)(gcol & 1);
        const uint64_t stamp = ((gcol >> 1) & 1) ? STAMP : 0ull;
        FSlot* sl = P.slots + (int64_t)pb * G;
        const int pub = ww >= 0 ? ww : 0;
        if (wv == pub) {
            if (ww >= 0) {
                // stores under a scalar branch per row: a select chain over the
                // row index would be turned into an indexed (scratch) access
                const int si = __builtin_amdgcn_readfirstlane(s_idx[ww]);
                T* crow = P.crow + ((int64_t)pb * G + g) * 2 * FB;
#pragma unroll
                for (int i = 0; i < RW; ++i)
                    if (i == si) {
                        cstore_t(crow + lane, pr[i]);
                        cstore_t(crow + FB + lane, zr[i]);
                    }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (lane == 0) {
                st_agent(&sl[g].val, (ww >= 0 ? (uint64_t)__double_as_longlong(wb) : 0ull) | stamp);
                st_agent(&sl[g].row, (ww >= 0 ? (uint64_t)wr : ROW_NONE) | stamp);
            }
        }
        const unsigned long long t1b = wall_clock64();
        // 3. wave 0 waits for every workgroup's slot of this column and picks
        //    the winner (max |x|, smallest row on ties); bounded wait
        if (wv == 0) {
            double gb = -1.0;
            int64_t gr = -1;
            int gq = -1;
            for (int64_t it = 0;; ++it) {
                bool ready = true;
                double lb = -1.0;
                int64_t lr = -1;
                int lq = -1;
                for (int q = lane; q < G; q += 64) {
                    const uint64_t v = ld_agent(&sl[q].val), r = ld_agent(&sl[q].row);
                    if ((v & STAMP) != stamp || (r & STAMP) != stamp) {
                        ready = false;
                        continue;
                    }
                    const uint64_t rr = r & ~STAMP;
                    if (rr == ROW_NONE) continue;
                    const double dv = __longlong_as_double((long long)(v & ~STAMP));
                    if (cand_better(dv, (int64_t)rr, lb, lr)) {
                        lb = dv;
                        lr = (int64_t)rr;
                        lq = q;
                    }
                }
                if (__all(ready)) {
                    gb = lb;
                    gr = lr;
                    gq = lq;
                    break;
                }
                if (it > P.spin_limit ||
                    ((it & 1023) == 1023 && __hip_atomic_load(P.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    if (lane == 0) __hip_atomic_store(P.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;  // gr stays -1: no pivot; the call reports the error
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double v = __shfl_xor(gb, off);
                const int64_t r = __shfl_xor(gr, off);
                const int q = __shfl_xor(gq, off);
                if (cand_better(v, r, gb, gr)) {
                    gb = v;
                    gr = r;
                    gq = q;
                }
            }
            if (lane == 0) {
                s_win_val = gb;
                s_win_row = gr;
                s_win_q = gq;
            }
        }
        __syncthreads();
        const double gb = s_win_val;
        const int64_t gr = s_win_row;
        const int gq = s_win_q;
        const unsigned long long t2 = wall_clock64();
        if (!(gr >= 0 && gb > P.tol)) {  // no pivot in this column (same decision everywhere)
            if (g == 0 && threadIdx.x == 0) {
                P.tphase[0] += t1 - t0;
                P.tphase[1] += t1b - t1;
                P.tphase[2] += t2 - t1b;
                P.tphase[5] += 1;
            }
            continue;
        }
        // 4. the pivot row; its own slot k of Z becomes 1
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const T* wrow = P.crow + ((int64_t)pb * G + gq) * 2 * FB;
        const T pvP = cload_t(wrow + lane, Ops::zero());
        T pvZ = cload_t(wrow + FB + lane, Ops::zero());
        pvZ = Ops::sel(lane == k, Ops::one(), pvZ);
        const double hsync = Ops::abs(bcast(pvP, c));  // waits for the loads (timer below)
        const unsigned long long t2b = wall_clock64() + (hsync < 0.0 ? 1ull : 0ull);
        const T hinv = Ops::inv(bcast(pvP, c));  // one division per column, products per row
        // 5. reduce every other row (pivot rows of earlier columns too) and
        //    normalise the pivot row; branch-free per row
        const T pvPn = Ops::mul(pvP, hinv), pvZn = Ops::mul(pvZ, hinv);
        const int64_t ploc = gr - row0;  // the pivot's index among this wave's rows, if it is one of them
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const bool isp = ploc == i;
            const T f = Ops::mul(bcast(pr[i], c), hinv);
            T np = Ops::fms(pr[i], f, pvP), nz = Ops::fms(zr[i], f, pvZ);
            np = Ops::sel(isp, pvPn, np);
            nz = Ops::sel(isp, pvZn, nz);
            np = Ops::sel(lane == c, Ops::sel(isp, Ops::one(), Ops::zero()), np);
            pr[i] = np;
            zr[i] = nz;
        }
        if (ploc >= 0 && ploc < RW) {
            live &= ~(1u << ploc);
            pmask |= 1u << ploc;
            if (lane == 0) s_slot[wv * RW + ploc] = k;
        }
        if (threadIdx.x == 0) s_pcol[k] = c;
        if (g == 0 && threadIdx.x == 0) {
            P.piv_rows[rank0 + k] = gr;
            P.piv_cols[rank0 + k] = P.J + c;
            P.qrows[k] = gr;
            const unsigned long long t3 = wall_clock64();
            P.tphase[0] += t1 - t0;
            P.tphase[1] += t1b - t1;
            P.tphase[2] += t2 - t1b;
            P.tphase[3] += t2b - t2;
            P.tphase[4] += t3 - t2b;
            P.tphase[5] += 1;
        }
        ++k;
    }
    __syncthreads();
    const int pc_lane = lane < k ? s_pcol[lane] : 0;
    // multipliers -P[r, pc] from the panel as it was (A still holds it),
    // then the reduced panel columns and Z[Q]
#pragma unroll
    for (int i = 0; i < RW; ++i) {
        if (!((valid >> i) & 1u)) continue;
        const int64_t r = row0 + i;
        const T orig = Ops::sel(lane < P.w, P.A[r * P.ld + P.J + lc], Ops::zero());
        const T mult = shfl_t(orig, pc_lane);  // lane t: row r's entry in pivot column pc[t]
        if (lane < P.w) P.A[r * P.ld + P.J + lane] = pr[i];
        const bool piv = (pmask >> i) & 1u;
        P.ZL[r * FB + lane] = Ops::sel(piv || lane >= k, Ops::zero(), Ops::neg(mult));
        if (piv) {
            P.ZQ[(int64_t)s_slot[wv * RW + i] * FB + lane] = zr[i];
            if (lane == 0) {
                P.used[r] = 1;
                P.pepoch[r] = P.panel;
            }
        }
    }
    if (g == 0 && threadIdx.x == 0) {
        *P.kpanel = k;
        *P.rank = rank0 + k;
    }
}

__global__ void k_fpanel_init(uint64_t* slots, int64_t nwords, int* err, int32_t* pepoch, int64_t N,
                              unsigned long long* tphase, int reset_timers) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x == 0) *err = 0;
    if (reset_timers && x < 8) tphase[x] = 0;
    // stamp bit set: differs from the stamp of the first two columns
    for (int64_t i = x; i < nwords; i += (int64_t)gridDim.x * blockDim.x) slots[i] = STAMP;
    for (int64_t r = x; r < N; r += (int64_t)gridDim.x * blockDim.x) pepoch[r] = -1;
}

// Right operand of the trailing update as real doubles, Kpad x nTr with
// Kpad = round_up(kReal * k, 16): f64: BQ[t][n] = A[Q[t], T0 + n];
// complex: rows 2t, 2t+1 = [Br, Bi], [-Bi, Br] interleaved per column.
template <class T>
__global__ __launch_bounds__(256) void k_fgather(const T* __restrict__ A, int64_t ld, int64_t T0, int64_t nT,
                                                 const int64_t* __restrict__ qrows, const int32_t* __restrict__ kpanel,
                                                 double* __restrict__ BQ) {
    constexpr int R = FloatOps<T>::kReal;
    const int k = *kpanel;
    const int Kpad = (int)round_up((int64_t)R * k, 16);
    const int t = blockIdx.y;  // pivot slot
    if (R * t >= Kpad) return;
    const int64_t nTr = R * nT;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nT; n += (int64_t)gridDim.x * blockDim.x) {
        if constexpr (R == 1) {
            BQ[(int64_t)t * nTr + n] = t < k ? A[qrows[t] * ld + T0 + n] : 0.0;
        } else {
            CplxF b{0.0, 0.0};
            if (t < k) b = A[qrows[t] * ld + T0 + n];
            double* r0 = BQ + (int64_t)(2 * t) * nTr + 2 * n;
            double* r1 = r0 + nTr;
            r0[0] = b.re;
            r0[1] = b.im;
            r1[0] = -b.im;
            r1[1] = b.re;
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 matrix-core GEMM for the trailing update (real view):
//   pivot tiles: C[Q[t], :] = LQ[t, :K] B[:K, :] for slots t < k   (B = A[Q, T]),
//   main tiles:  C[r, :] += L[r, :K] B[:K, :]   for rows r that are not pivots
//                of this panel (pepoch[r] != panel)             (B = Y, L = -P[:, pc]).
// The two kinds run as separate launches, pivot tiles first.
// 64 x 64 tile per 256-thread workgroup (4 waves of 32 x 32 = 2 x 2 MFMA
// 16x16x4 blocks), K staged through LDS in chunks of 16, double-buffered.
// ---------------------------------------------------------------------------

constexpr int FG_M = 64, FG_N = 64, FG_KC = 16, FG_T = 256;
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

struct FGemmArgs {
    const double* L;   // [N][ldl]
    const double* LQ;  // [FB][ldl]
    int64_t ldl;
    const double* B;   // [Kpad][nTr]
    double* C;         // A (real view) + real offset of T0
    int64_t ldc;
    int64_t N, nTr;
    const int32_t* kpanel;
    int kreal;         // doubles per element (1 or 2)
    const int32_t* pepoch;
    int panel;
    const int64_t* qrows;
    int64_t tiles_q;   // pivot tiles come first in the grid
    int64_t tiles_n;   // column tiles
    int64_t tiles_main;
};

// F32: the operands are rounded to FP32 in LDS, the products accumulated by
// FP32 matrix cores and the sums added to the FP64 matrix (the mixed method).
template <bool F32>
__global__ __launch_bounds__(FG_T, 2) void k_fgemm(FGemmArgs a) {
    using E = typename std::conditional<F32, float, double>::type;
    using V = typename std::conditional<F32, v4f, v4d>::type;
    __shared__ E sA[2][FG_M][FG_KC + 1];
    __shared__ E sB[2][FG_KC][FG_N + 1];
    const int k = *a.kpanel;
    if (k == 0) return;
    const int K = a.kreal * k;
    const int nchunks = (int)ceil_div(K, FG_KC);
    int64_t bid = blockIdx.x;
    const bool qtile = bid < a.tiles_q;
    int64_t tm, tn;
    if (qtile) {
        tm = 0;
        tn = bid;
    } else {
        // XCD-aware: consecutive workgroup ids land on different XCDs; give each
        // XCD a contiguous range of tiles, walked in groups of 8 tile rows
        const int64_t total = (int64_t)gridDim.x - a.tiles_q;  // a multiple of 8
        const int64_t b = bid - a.tiles_q;
        const int64_t per = total / 8;
        const int64_t t = (b % 8) * per + b / 8;
        if (t >= a.tiles_main) return;
        const int64_t mt = ceil_div(a.N, FG_M);
        const int64_t gsz = 8 * a.tiles_n;
        const int64_t gi = t / gsz, tt = t - gi * gsz;
        const int64_t gm = min((int64_t)8, mt - gi * 8);
        tm = gi * 8 + tt % gm;
        tn = tt / gm;
    }
    const int64_t m0 = tm * FG_M, n0 = tn * FG_N;
    const int64_t Mrows = qtile ? k : a.N;
    const double* L = qtile ? a.LQ : a.L;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    // loaders: A chunk 64 rows x 16 = 4 doubles per thread; B chunk 16 x 64 = 4 per thread
    const int ar = tid >> 2, ak = (tid & 3) * 4;
    const int bk = tid >> 4, bn = (tid & 15) * 4;
    double ra[4], rb[4];
    auto load = [&](int ch) {
        const int kc = ch * FG_KC;
        const int64_t grow = m0 + ar;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int kk = kc + ak + e;
            ra[e] = (grow < Mrows && kk < K) ? L[grow * a.ldl + kk] : 0.0;
        }
        const int kb = kc + bk;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t col = n0 + bn + e;
            rb[e] = (kb < K && col < a.nTr) ? a.B[(int64_t)kb * a.nTr + col] : 0.0;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int e = 0; e < 4; ++e) sA[buf][ar][ak + e] = (E)ra[e];
#pragma unroll
        for (int e = 0; e < 4; ++e) sB[buf][bk][bn + e] = (E)rb[e];
    };
    V acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = V{0, 0, 0, 0};
    load(0);
    store(0);
    __syncthreads();
    const int r16 = lane & 15, kg = lane >> 4;
    for (int ch = 0; ch < nchunks; ++ch) {
        const int buf = ch & 1;
        if (ch + 1 < nchunks) load(ch + 1);
#pragma unroll
        for (int ks = 0; ks < FG_KC / 4; ++ks) {
            E av[2], bv[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) av[mi] = sA[buf][wm * 32 + mi * 16 + r16][ks * 4 + kg];
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) bv[ni] = sB[buf][ks * 4 + kg][wn * 32 + ni * 16 + r16];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
                    if constexpr (F32)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
                    else
                        acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
                }
        }
        if (ch + 1 < nchunks) {
            store(buf ^ 1);
            __syncthreads();
        }
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            // v_mfma_f64_16x16x4f64 output: lane holds D[lane / 16 + 4 v][lane % 16]
            // (verified on gfx950); the FP32 16x16x4 form, like the int8
            // 16x16x64 one, D[4 (lane / 16) + v][lane % 16]
            const int64_t row = m0 + wm * 32 + mi * 16 + (F32 ? 4 * kg + v : kg + 4 * v);
            if (row >= Mrows) continue;
            double* crow;
            if (qtile) {
                crow = a.C + a.qrows[row] * a.ldc;
            } else {
                if (a.pepoch[row] == a.panel) continue;
                crow = a.C + row * a.ldc;
            }
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const int64_t col = n0 + wn * 32 + ni * 16 + r16;
                if (col >= a.nTr) continue;
                if (qtile) crow[col] = (double)acc[mi][ni][v];
                else crow[col] += (double)acc[mi][ni][v];
            }
        }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

struct FloatBlockedLayout {
    size_t ZL, ZQ, qrows, kpanel, pepoch, slots, crow, err, tphase, BQ, total;
};

template <class T>
static FloatBlockedLayout float_layout(int64_t N, int64_t M, int64_t G) {
    auto al = [](size_t x) { return (size_t)round_up((int64_t)x, 256); };
    FloatBlockedLayout l{};
    size_t off = 0;
    l.ZL = off; off += al((size_t)N * FB * sizeof(T));
    l.ZQ = off; off += al((size_t)FB * FB * sizeof(T));
    l.qrows = off; off += al(FB * 8);
    l.kpanel = off; off += al((size_t)4 * ceil_div(std::max<int64_t>(M, 1), FB));
    l.pepoch = off; off += al((size_t)N * 4);
    l.slots = off; off += al((size_t)2 * G * sizeof(FSlot));
    l.crow = off; off += al((size_t)2 * G * 2 * FB * sizeof(T));
    l.err = off; off += 256;
    l.tphase = off; off += 256;
    l.BQ = off; off += al((size_t)(FloatOps<T>::kReal * FB + 16) * FloatOps<T>::kReal * std::max<int64_t>(M, 1) * 8);
    l.total = off;
    return l;
}

// Dynamic LDS that keeps one panel workgroup per CU when the grid fits the
// CUs (the column loop is latency-bound: workgroups sharing a CU's SIMDs
// slow every column down); 0 when more workgroups than CUs are needed.
constexpr unsigned FPANEL_SPREAD_LDS = 96 * 1024;

template <class T, int RW>
static int cooperative_fits(int64_t G, DeviceContext* ctx, bool* fits, unsigned* lds) {
    int cus = 0;
    LINAC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    *lds = G <= cus ? FPANEL_SPREAD_LDS : 0u;
    int per_cu = 0;
    LINAC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fpanel<T, RW>, 256, *lds));
    *fits = (int64_t)per_cu * cus >= G;
    return LINAC_OK;
}

// One blocked float sweep, resumable panel by panel (FloatSweep, declared in
// float_blocked.hpp): set-up, any number of run() calls, finish().
template <class T>
int FloatSweep<T>::setup(T* A, int64_t ld, int64_t N, int64_t M, double tol, hipStream_t s, int64_t* rank,
                         uint8_t* used, int64_t* piv_rows, int64_t* piv_cols, DeviceContext* ctx, ScratchSlot slot,
                         FloatFactors* keep_, void* ws) {
    bool fits = false;
    // rows per wave: 16 where that grid is resident, else 32.  For float64
    // the knob fpanel_rows (per size class) can ask for 32: half the
    // workgroups exchange candidates per column, each updating twice the
    // rows (faster from about N = 6000, slower below 3000; complex128, with
    // twice the registers per row, keeps 16)
    int rw = 16;
    const bool want32 =
        std::is_same<T, double>::value && tuning().pick(Tuning::FPANEL_ROWS, std::max(N, M)) == 32;
    if (!want32) LINAC_TRY((cooperative_fits<T, 16>(ceil_div(N, FWAVES * 16), ctx, &fits, &lds)));
    if (!fits) {
        LINAC_TRY((cooperative_fits<T, 32>(ceil_div(N, FWAVES * 32), ctx, &fits, &lds)));
        rw = 32;
    }
    if (!fits) return 1;
    kfn = rw == 16 ? (const void*)k_fpanel<T, 16> : (const void*)k_fpanel<T, 32>;
    G = ceil_div(N, (int64_t)FWAVES * rw);
    this->ctx = ctx;
    keep = keep_;
    const FloatBlockedLayout l = float_layout<T>(N, M, G);
    void* base = ws;
    if (!base) LINAC_TRY(ctx->get(slot, l.total, &base));
    char* b = (char*)base;
    npanels = ceil_div(M, FB);
    next = 0;
    // with `keep`, every panel's factors (ZL, ZQ, pivot rows) get their own
    // slice, kept after the sweep for the refinement's replay
    if (keep) {
        const size_t zl = (size_t)round_up(N * FB * (int64_t)sizeof(T), 256);
        const size_t zq = (size_t)FB * FB * sizeof(T), q = (size_t)FB * 8;
        void* kb = nullptr;
        LINAC_TRY(ctx->get(SLOT_AUX2, (size_t)npanels * (zl + zq + q), &kb));
        keep->ZL = kb;
        keep->ZQ = (char*)kb + npanels * zl;
        keep->qrows = (int64_t*)((char*)kb + npanels * (zl + zq));
        keep->npanels = npanels;
        keep->pepoch = (const int32_t*)(b + l.pepoch);
    }
    P = FPanelArgs<T>{};
    P.A = A;
    P.ld = ld;
    P.N = N;
    P.M = M;
    P.tol = tol;
    P.ZL = (T*)(b + l.ZL);
    P.ZQ = (T*)(b + l.ZQ);
    P.qrows = (int64_t*)(b + l.qrows);
    kpanel = (int32_t*)(b + l.kpanel);
    P.pepoch = (int32_t*)(b + l.pepoch);
    P.used = used;
    P.piv_rows = piv_rows;
    P.piv_cols = piv_cols;
    P.rank = rank;
    P.slots = (FSlot*)(b + l.slots);
    P.crow = (T*)(b + l.crow);
    P.err = (int*)(b + l.err);
    P.tphase = (unsigned long long*)(b + l.tphase);
    P.spin_limit = (int64_t)1 << 24;  // ~ seconds of polling: far beyond any healthy barrier
    BQ = (double*)(b + l.BQ);
    k_fpanel_init<<<(unsigned)std::min<int64_t>(1024, ceil_div(std::max<int64_t>(N, 4 * G), 256)), 256, 0, s>>>(
        (uint64_t*)P.slots, 4 * G, P.err, P.pepoch, N, P.tphase, 1);
    LINAC_HIP(hipGetLastError());
    return LINAC_OK;
}

template <class T>
size_t FloatSweep<T>::workspace_bytes(int64_t N, int64_t M) {
    // the grid is largest with 16 rows per wave
    return float_layout<T>(N, M, ceil_div(N, (int64_t)FWAVES * 16)).total;
}

template <class T>
int FloatSweep<T>::run(hipStream_t s, int64_t max_panels, Profiler* prof) {
    constexpr int R = FloatOps<T>::kReal;
    const int64_t N = P.N, M = P.M, ld = P.ld;
    T* A = P.A;
    const int64_t stop = max_panels < 0 ? npanels : std::min(npanels, next + max_panels);
    for (; next < stop; ++next) {
        const int64_t pi = next;
        P.J = pi * FB;
        P.w = (int)std::min<int64_t>(FB, M - P.J);
        P.panel = (int)pi;
        P.kpanel = kpanel + pi;
        if (keep) {
            P.ZL = (T*)keep->ZL + pi * (round_up(N * FB * (int64_t)sizeof(T), 256) / (int64_t)sizeof(T));
            P.ZQ = (T*)keep->ZQ + pi * FB * FB;
            P.qrows = keep->qrows + pi * FB;
        }
        void* args[] = {&P};
        LINAC_HIP(hipLaunchCooperativeKernel(kfn, dim3((unsigned)G), dim3(256), args, lds, s));
        const int64_t T0 = P.J + P.w, nT = M - T0;
        if (nT <= 0) continue;
        const int64_t nTr = R * nT;
        const dim3 ggrid((unsigned)std::min<int64_t>(256, ceil_div(nT, 256)), FB + 16 / R);
        k_fgather<T><<<ggrid, 256, 0, s>>>(A, ld, T0, nT, P.qrows, P.kpanel, BQ);
        FGemmArgs ga{};
        ga.L = (const double*)P.ZL;
        ga.LQ = (const double*)P.ZQ;
        ga.ldl = (int64_t)R * FB;
        ga.B = BQ;
        ga.C = (double*)(A + T0);
        ga.ldc = R * ld;
        ga.N = N;
        ga.nTr = nTr;
        ga.kpanel = P.kpanel;
        ga.kreal = R;
        ga.pepoch = P.pepoch;
        ga.panel = P.panel;
        ga.qrows = P.qrows;
        ga.tiles_n = ceil_div(nTr, FG_N);
        // 1. Y = Z[Q] A[Q, T] into the pivot rows (one tile row: k <= 64)
        ga.tiles_q = ga.tiles_n;
        ga.tiles_main = 0;
        k_fgemm<false><<<(unsigned)ga.tiles_q, FG_T, 0, s>>>(ga);
        // 2. the right operand becomes Y; every other row: A[r, T] -= P[r, pc] Y
        k_fgather<T><<<ggrid, 256, 0, s>>>(A, ld, T0, nT, P.qrows, P.kpanel, BQ);
        ga.tiles_q = 0;
        ga.tiles_main = ceil_div(N, FG_M) * ga.tiles_n;
        // algorithmic work of the launch: N x nTr x (kReal * k) real MACs; k is
        // known after the sweep, so the sample holds N x nTr until then
        if (prof) LINAC_TRY(prof->sample_begin(s, pi, 16.0 * (double)N * (double)nTr, (double)N * (double)nTr));
        if (f32_update) k_fgemm<true><<<(unsigned)round_up(ga.tiles_main, 8), FG_T, 0, s>>>(ga);
        else k_fgemm<false><<<(unsigned)round_up(ga.tiles_main, 8), FG_T, 0, s>>>(ga);
        if (prof) LINAC_TRY(prof->sample_end(s));
        LINAC_HIP(hipGetLastError());
    }
    return LINAC_OK;
}

template <class T>
int FloatSweep<T>::finish(hipStream_t s, Profiler* prof) {
    constexpr int R = FloatOps<T>::kReal;
    int err = 0;
    std::vector<int32_t> kp((size_t)npanels);
    LINAC_HIP(hipMemcpyAsync(&err, P.err, sizeof(int), hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipMemcpyAsync(kp.data(), kpanel, kp.size() * 4, hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipStreamSynchronize(s));
    if (err) return fail(LINAC_ERR_DEVICE, "blocked float elimination: a pivot exchange between workgroups timed out");
    if (getenv("LINAC_DEBUG_TIMING")) {
        unsigned long long tp[8];
        int rate_khz = 0;
        if (hipMemcpy(tp, P.tphase, sizeof(tp), hipMemcpyDeviceToHost) == hipSuccess && tp[5] &&
            hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess) {
            const double us = 1e3 / (double)rate_khz;
            fprintf(stderr,
                    "k_fpanel phases over %llu columns (us/column): search %.2f publish %.2f wait %.2f "
                    "pivot-row %.2f update %.2f\n",
                    tp[5], tp[0] * us / tp[5], tp[1] * us / tp[5], tp[2] * us / tp[5], tp[3] * us / tp[5],
                    tp[4] * us / tp[5]);
        }
    }
    if (prof)
        for (auto& x : prof->samples) x.macs *= (double)R * kp[(size_t)x.tag];
    if (keep) keep->kp = kp;
    return LINAC_OK;
}

// Returns 1 (not applicable) when the cooperative panel grid cannot be resident.
template <class T>
int float_blocked_run(T* A, int64_t ld, int64_t N, int64_t M, double tol, hipStream_t s, int64_t* rank,
                      uint8_t* used, int64_t* piv_rows, int64_t* piv_cols, DeviceContext* ctx, Profiler* prof,
                      FloatFactors* keep, bool f32_update) {
    FloatSweep<T> sw;
    const int rc = sw.setup(A, ld, N, M, tol, s, rank, used, piv_rows, piv_cols, ctx, SLOT_PANEL_A, keep);
    if (rc != LINAC_OK) return rc;
    sw.f32_update = f32_update;
    LINAC_TRY(sw.run(s, -1, prof));
    return sw.finish(s, prof);
}

template struct FloatSweep<double>;
template struct FloatSweep<CplxF>;
template int float_blocked_run<double>(double*, int64_t, int64_t, int64_t, double, hipStream_t, int64_t*, uint8_t*,
                                       int64_t*, int64_t*, DeviceContext*, Profiler*, FloatFactors*, bool);
template int float_blocked_run<CplxF>(CplxF*, int64_t, int64_t, int64_t, double, hipStream_t, int64_t*, uint8_t*,
                                      int64_t*, int64_t*, DeviceContext*, Profiler*, FloatFactors*, bool);

}  // namespace linac
