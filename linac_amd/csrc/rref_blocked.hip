// Blocked right-looking Gauss-Jordan over GF(p), p <= 4278124287 (uint32).
//
// The column range is swept in outer panels of BO = 256 columns.  For each
// outer panel the library keeps a panel buffer X = [P | Z] (N x 512, u32):
//   P = A[:, J:J+BO]   the panel columns of every row,
//   Z                  the accumulated row transformation restricted to the
//                      panel's pivot rows Q ("augmented identity": when row r
//                      becomes the pivot of slot q, Z[r, q] is set to 1, and
//                      every later row operation is applied to [P | Z]).
// Inner steps eliminate BI = 32 columns at a time inside X:
//   1. candidates S: the first CI = 32 unused rows with a non-zero entry in the
//      inner columns (k_row_flags + ordered compaction);
//   2. one workgroup runs fraction-free Gauss-Jordan on [X[S, cols] | I] in LDS,
//      giving the pivot rows R, their columns, and Apiv^{-1} (the identity part
//      restricted to R, since pivot rows only ever absorb other pivot rows);
//   3. U = Apiv^{-1} X[R, c0:]  (normalised pivot rows),
//      X[r, c0:] -= X[r, c0 + Cp] U  for r not in R  (int8-limb MFMA GEMM),
//      X[R, c0:] = U.
// If the candidates did not exhaust the rows with non-zero inner entries and
// fewer than BI pivots were found, the same inner panel is processed again
// (its remaining rows have been reduced by the pivots just found, so every
// repeat adds a pivot).  Inner panels must complete in column order: a row
// whose leading entry lies in an unfinished panel must not become the pivot of
// a later column.  The panel position lives in device memory, so the host
// launches a fixed sequence of step groups and synchronises once per outer
// panel to see whether extra groups are needed.
// After the inner steps Z holds Phi[:, Q] of the panel's net transformation
// Phi, and the trailing columns T = [J+BO, M) get, in two MFMA GEMMs,
//   A[Q, T] <- Z[Q] A[Q, T],      A[r, T] <- A[r, T] + Z[r] A[Q, T]  (r not in Q).
// The dominant cost is the second GEMM: N x |Q| x |T| modular multiply-adds
// with K = |Q| <= 256.
//
// Rows are never swapped; `used` marks pivot rows, the pivot list records
// (row, column) and the shared finalisation gathers rows in column order.
// Pivot rows come out normalised.  Replaces the column loop of
// linac/pycuda_row_reduce.py:94-131 for large GF(p) matrices.
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <type_traits>

#include "gf_common.hpp"
#include "modp_gemm.hpp"

namespace linac {

constexpr int BO = 256;        // outer panel width (GEMM depth)
constexpr int BI = 32;         // inner panel width
constexpr int CI = 32;         // candidate rows per inner step
constexpr int XLD = 2 * BO;    // leading dimension of the panel buffer [P | Z]
constexpr int KPI = 64;        // padded depth of the inner GEMM

struct BlockState {
    int32_t k;           // pivots found by the last inner step
    int32_t q0;          // pivot slots used in the current outer panel
    int32_t c0_next;     // first column (in the panel) of the next inner step
    int32_t c0_cur;      // first column of the last inner step
    int32_t bcur;        // width of the current outer panel
    int32_t pad_;
    int64_t R[CI];       // pivot rows of the last inner step (pivot order)
    int32_t Cp[CI];      // their columns relative to the inner panel start
    uint32_t apinv[CI * CI];  // Montgomery form of (X[R, c0 + Cp])^{-1}
    unsigned long long tphase[8];  // accumulated wall-clock ticks per phase of k_inner_factor
};

// ---------------------------------------------------------------------------
// panel buffer management
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_panel_load(uint32_t* __restrict__ X, const uint32_t* __restrict__ A,
                                                    int64_t ld, int64_t N, int64_t J, int Bcur) {
    const int64_t total = N * XLD;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = x / XLD;
        const int j = (int)(x - r * XLD);
        X[x] = j < Bcur ? A[r * ld + J + j] : 0u;
    }
}

__global__ __launch_bounds__(256) void k_panel_store(uint32_t* __restrict__ A, const uint32_t* __restrict__ X,
                                                     int64_t ld, int64_t N, int64_t J, int Bcur) {
    const int64_t total = N * Bcur;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = x / Bcur;
        const int j = (int)(x - r * Bcur);
        A[r * ld + J + j] = X[r * XLD + j];
    }
}

__global__ __launch_bounds__(256) void k_epoch_init(int32_t* epoch, int64_t N) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x)
        epoch[r] = -1;
}

__global__ void k_block_state_init(BlockState* bs, int bcur, int reset_timers) {
    if (reset_timers)
        for (int i = 0; i < 8; ++i) bs->tphase[i] = 0;
    bs->k = 0;
    bs->q0 = 0;
    bs->c0_next = 0;
    bs->c0_cur = 0;
    bs->bcur = bcur;
}

// ---------------------------------------------------------------------------
// inner factorisation: one workgroup
// ---------------------------------------------------------------------------

constexpr int IF_T = 1024;  // threads of the inner factorisation: 16 waves, two candidate rows each
static_assert(IF_T / 64 * 2 == CI, "each wave owns two candidate rows");
static_assert(BI + CI == 64, "one augmented column per lane");

// One workgroup.  Candidate rows live in LDS as [panel columns | identity];
// wave w owns rows 2w and 2w+1.  Per column every wave ballots the column
// itself (all waves see the same data, so they agree on the pivot without a
// barrier), then updates its own rows with lane j handling column j.  Within
// a wave the reads of a row precede its writes in program order, so one
// barrier per column suffices.  No rows are swapped: pivots are recorded by
// their LDS row, and the identity columns index LDS rows directly.
__global__ __launch_bounds__(IF_T) void k_inner_factor(uint32_t* __restrict__ X, int64_t N,
                                                      uint8_t* __restrict__ used, int32_t* __restrict__ slot_of,
                                                      int32_t* __restrict__ epoch_of, int epoch,
                                                      int64_t colbase, FieldU32 f, ElimState* st,
                                                      int64_t* __restrict__ piv_rows, int64_t* __restrict__ piv_cols,
                                                      BlockState* bs) {
    constexpr int W2 = BI + CI;
    __shared__ uint32_t aug[CI][W2 + 1];
    __shared__ int64_t cand[CI + 1];
    __shared__ int prow_of[CI];   // LDS row of the t-th pivot
    __shared__ int pcol[CI];      // its column within the inner panel
    __shared__ int s_total;
    __shared__ int s_wcnt[IF_T / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const unsigned long long t0 = wall_clock64();
    const int c0 = bs->c0_next;
    const int bcur = bs->bcur;
    if (c0 >= bcur) {  // every inner panel of this outer panel is complete
        if (tid == 0) bs->k = 0;
        return;
    }
    const int w = min(BI, bcur - c0);

    if (tid == 0) s_total = 0;
    __syncthreads();
    // 1. the first CI (+1 to know whether more exist) candidate rows, in row order
    for (int64_t base = st->first_unused; base < N; base += IF_T) {
        const int64_t r = base + tid;
        bool fl = false;
        if (r < N && !used[r]) {  // unused row with a non-zero entry in the inner panel
            const uint4* row = reinterpret_cast<const uint4*>(X + r * XLD + c0);  // 128-byte aligned
            uint4 v[BI / 4];
#pragma unroll
            for (int q = 0; q < BI / 4; ++q) v[q] = row[q];
            uint32_t acc = 0;
#pragma unroll
            for (int q = 0; q < BI / 4; ++q) {
                const int j = 4 * q;
                acc |= (j + 0 < w ? v[q].x : 0u) | (j + 1 < w ? v[q].y : 0u) | (j + 2 < w ? v[q].z : 0u) |
                       (j + 3 < w ? v[q].w : 0u);
            }
            fl = acc != 0u;
        }
        const unsigned long long m = __ballot(fl);
        if (lane == 0) s_wcnt[wid] = __popcll(m);
        __syncthreads();
        int before = s_total;
        for (int q = 0; q < wid; ++q) before += s_wcnt[q];
        const int pos = before + __popcll(m & ((1ull << lane) - 1ull));
        if (fl && pos <= CI) cand[pos] = r;
        __syncthreads();
        if (tid == 0) {
            int sum = 0;
            for (int q = 0; q < IF_T / 64; ++q) sum += s_wcnt[q];
            s_total += sum;
        }
        __syncthreads();
        if (s_total > CI) break;
    }
    const unsigned long long t1 = wall_clock64();
    const int total = s_total;
    const int nC = total < CI ? total : CI;
    const int more = total > CI ? 1 : 0;
    if (nC == 0) {
        if (tid == 0) {
            bs->k = 0;
            bs->c0_cur = c0;
            bs->c0_next = c0 + BI;
        }
        return;
    }

    // 2. [X[S, c0:c0+w] | I] in LDS; rows >= nC are zero and never pivot
    {
        const int i = tid >> 6, j = lane;  // two passes of 16 rows
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int ii = i + 16 * half;
            uint32_t v = 0u;
            if (ii < nC) {
                if (j < w) v = X[cand[ii] * XLD + c0 + j];
                else if (j >= BI) v = (j - BI == ii) ? 1u : 0u;
            }
            aug[ii][j] = v;
        }
    }
    __syncthreads();

    const unsigned long long t2 = wall_clock64();
    // 3. fraction-free Gauss-Jordan: row_i <- (h * row_i - a_i * row_pv) R^{-1}
    //    (rows stay non-zero multiples of the exact Gauss-Jordan rows)
    unsigned long long pmask = 0ull;  // LDS rows already used as pivots (same in every wave)
    int t = 0;
    const int r0 = 2 * wid, r1 = 2 * wid + 1;
    for (int c = 0; c < w && t < nC; ++c) {
        const bool nz = lane < nC && !((pmask >> lane) & 1ull) && aug[lane < CI ? lane : 0][c] != 0u;
        const unsigned long long m = __ballot(nz);
        if (m == 0ull) continue;  // no pivot in this column (every wave sees the same)
        const int pv = __ffsll((long long)m) - 1;
        const uint32_t h = aug[pv][c];
        const uint32_t y = aug[pv][lane];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int r = q ? r1 : r0;
            if (r < nC && r != pv) {
                const uint32_t a = aug[r][c];
                const uint32_t x = aug[r][lane];
                aug[r][lane] = f.ff_update(h, x, a, y);
            }
        }
        if (tid == 0) {
            prow_of[t] = pv;
            pcol[t] = c;
        }
        pmask |= 1ull << pv;
        ++t;
        __syncthreads();
    }
    const int k = t;
    __syncthreads();
    const unsigned long long t3 = wall_clock64();

    // 4. normalise the pivot rows.  The k inverses run side by side in the
    //    lanes of wave 0 (one Fermat chain in total instead of one per wave,
    //    which would serialise on the SIMDs), then each wave scales its rows.
    __shared__ uint32_t s_invm[CI];
    __shared__ int s_tof[CI];
    if (wid == 0 && lane < k) {
        const int r = prow_of[lane];
        s_invm[lane] = f.to_mont(f.inverse(aug[r][pcol[lane]]));
        s_tof[r] = lane;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = q ? r1 : r0;
        if ((pmask >> r) & 1ull) aug[r][lane] = f.mont(s_invm[s_tof[r]], aug[r][lane]);
    }
    __syncthreads();

    const unsigned long long t4 = wall_clock64();
    // 5. outputs: Apiv^{-1} (Montgomery form), pivot rows/columns, bookkeeping
    const int64_t rank0 = st->rank;
    const int q0 = bs->q0;
    for (int x = tid; x < k * k; x += IF_T) {
        const int a = x / k, b = x - a * k;
        bs->apinv[a * CI + b] = f.to_mont(aug[prow_of[a]][BI + prow_of[b]]);
    }
    if (tid < k) {
        const int64_t r = cand[prow_of[tid]];
        bs->R[tid] = r;
        bs->Cp[tid] = pcol[tid];
        used[r] = 1;
        epoch_of[r] = epoch;
        slot_of[r] = tid;
        piv_rows[rank0 + tid] = r;
        piv_cols[rank0 + tid] = colbase + c0 + pcol[tid];
        X[r * XLD + BO + q0 + tid] = 1u;  // Z[r, slot] = 1
    }
    __syncthreads();
    // advance first_unused past the rows that are now pivots, 1024 rows per probe
    __shared__ unsigned long long s_first;
    if (tid == 0) s_first = ~0ull;
    __syncthreads();
    for (int64_t base = st->first_unused; base < N; base += IF_T) {
        const int64_t r = base + tid;
        const unsigned long long m = __ballot(r < N && !used[r]);
        if (lane == 0 && m) atomicMin(&s_first, (unsigned long long)(base + wid * 64 + __ffsll((long long)m) - 1));
        __syncthreads();
        const bool found = s_first != ~0ull;
        __syncthreads();
        if (found) break;
    }
    if (tid == 0) {
        st->rank = rank0 + k;
        bs->k = k;
        bs->q0 = q0 + k;
        bs->c0_cur = c0;
        bs->c0_next = (more && k < w) ? c0 : c0 + BI;
        st->first_unused = s_first != ~0ull ? (int64_t)s_first : N;
        const unsigned long long t5 = wall_clock64();
        bs->tphase[0] += t1 - t0;
        bs->tphase[1] += t2 - t1;
        bs->tphase[2] += t3 - t2;
        bs->tphase[3] += t4 - t3;
        bs->tphase[4] += t5 - t4;
        bs->tphase[5] += 1;
    }
}

// U[t, n] = sum_t' Apiv^{-1}[t, t'] X[R[t'], n] for every column n of the panel
// buffer (columns left of the inner panel are zero in the pivot rows, so the
// product is zero there); writes U (u32) and its limb planes (B operand of
// the inner GEMM, planes[l][n][t], depth KPI).
__global__ __launch_bounds__(256) void k_make_u(const uint32_t* __restrict__ X, FieldU32 f,
                                                const BlockState* __restrict__ bs, uint32_t* __restrict__ U,
                                                int8_t* __restrict__ planes, int64_t plane_stride) {
    __shared__ uint32_t s_inv[CI * CI];
    __shared__ int64_t s_R[CI];
    const int k = bs->k;
    for (int x = threadIdx.x + threadIdx.y * 64; x < k * CI; x += 256) s_inv[x] = bs->apinv[x];
    if (threadIdx.y == 0 && threadIdx.x < k) s_R[threadIdx.x] = bs->R[threadIdx.x];
    __syncthreads();
    const int t = threadIdx.x;  // 0..63
    const int n = blockIdx.x * 4 + threadIdx.y;
    uint32_t u = 0;
    if (t < k) {
        for (int s = 0; s < k; ++s) {
            const uint32_t prod = f.mont(s_inv[t * CI + s], X[s_R[s] * XLD + n]);
            const uint32_t v = u + prod;
            u = v >= f.p ? v - f.p : v;
        }
        U[t * XLD + n] = u;
    }
    int8_t d[4] = {0, 0, 0, 0};
    if (t < k) balanced_digits(centred(u, f.p), d);
#pragma unroll
    for (int l = 0; l < 4; ++l) planes[l * plane_stride + (int64_t)n * KPI + t] = d[l];
}

// Limb planes of a left operand: planes[l][m][t] = digit l of (+/-) Y[row(m), col(t)]
// for t < K; zero for t in [K, Kpad) and for rows whose epoch lies in [e0, e1).
__global__ __launch_bounds__(256) void k_make_lhs(const uint32_t* __restrict__ Y, int64_t ldy, int64_t Mrows,
                                                  const int64_t* __restrict__ rowlist, int64_t col0_host,
                                                  const int32_t* col0ptr,
                                                  const int32_t* __restrict__ colidx, const int32_t* kptr, int Khost,
                                                  int Kpad, const int32_t* __restrict__ epoch_of, int e0, int e1,
                                                  int negate, uint32_t p, int8_t* __restrict__ planes,
                                                  int64_t plane_stride) {
    const int K = kptr ? *kptr : Khost;
    const int64_t col0 = col0ptr ? *col0ptr : col0_host;
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t m = x / Kpad;
    const int t = (int)(x - m * Kpad);
    if (m >= Mrows) return;
    const int64_t r = rowlist ? rowlist[m] : m;
    int8_t d[4] = {0, 0, 0, 0};
    const bool zero_row = epoch_of && epoch_of[r] >= e0 && epoch_of[r] < e1;
    if (t < K && !zero_row) {
        const int64_t c = col0 + (colidx ? colidx[t] : t);
        int32_t cv = centred(Y[r * ldy + c], p);
        balanced_digits(negate ? -cv : cv, d);
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) planes[l * plane_stride + m * Kpad + t] = d[l];
}

// Limb planes of a right operand, transposed: planes[l][n][t] = digit l of
// Y[rowlist[t], col0 + n] for t < K, zero for t in [K, Kpad).  64 x 64 tiles
// through LDS so that both the row reads and the plane writes are coalesced.
__global__ __launch_bounds__(256) void k_make_rhs(const uint32_t* __restrict__ Y, int64_t ldy,
                                                  const int64_t* __restrict__ rowlist, int K, int64_t col0,
                                                  int64_t Ncols, int Kpad, uint32_t p, int8_t* __restrict__ planes,
                                                  int64_t plane_stride) {
    __shared__ int32_t tile[64][65];
    const int64_t nb = (int64_t)blockIdx.x * 64;
    const int tb = blockIdx.y * 64;
    for (int x = threadIdx.x; x < 64 * 64; x += 256) {
        const int tt = x >> 6, nn = x & 63;
        const int t = tb + tt;
        const int64_t n = nb + nn;
        int32_t v = 0;
        if (t < K && n < Ncols) v = centred(Y[rowlist[t] * ldy + col0 + n], p);
        tile[tt][nn] = v;
    }
    __syncthreads();
    // each thread writes 16 consecutive t of one column n, for every limb
    const int nn = threadIdx.x >> 2, tq = (threadIdx.x & 3) * 16;
    const int64_t n = nb + nn;
    if (n >= Ncols) return;
    int8_t buf[4][16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        int8_t d[4];
        balanced_digits(tile[tq + i][nn], d);
#pragma unroll
        for (int l = 0; l < 4; ++l) buf[l][i] = d[l];
    }
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        v4i w;
        __builtin_memcpy(&w, buf[l], 16);
        *reinterpret_cast<v4i*>(planes + l * plane_stride + n * Kpad + tb + tq) = w;
    }
}

// Y[rowlist[t], col0 + n] = U[t, n] for t < K, n < ncols.
__global__ __launch_bounds__(256) void k_scatter_rows(uint32_t* __restrict__ Y, int64_t ldy,
                                                      const int64_t* __restrict__ rowlist, const int32_t* kptr,
                                                      int Khost, int64_t col0, int64_t ncols,
                                                      const uint32_t* __restrict__ U, int64_t ldu) {
    const int K = kptr ? *kptr : Khost;
    const int t = blockIdx.y;
    if (t >= K) return;
    const int64_t r = rowlist[t];
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < ncols; n += (int64_t)gridDim.x * blockDim.x)
        Y[r * ldy + col0 + n] = U[t * ldu + n];
}

// ---------------------------------------------------------------------------
// the modular GEMM: C <- (C - L B) mod p  (or C <- -L B mod p when CZERO)
// 256 threads, 64 x 64 tile per workgroup, 32 x 32 per wave (2 x 2 MFMA
// 16x16x64 tiles), seven diagonal accumulators per tile.
// ---------------------------------------------------------------------------

template <bool CZERO>
__global__ __launch_bounds__(256) void k_gemm_modp(uint32_t* __restrict__ C, int64_t ldc, int64_t Mrows,
                                                   int64_t Ncols, const int8_t* __restrict__ L,
                                                   int64_t planeL, const int8_t* __restrict__ B,
                                                   int64_t planeB, int Kpad, uint32_t p, double pd, double pinv,
                                                   const int32_t* kptr, const int32_t* lo_ptr,
                                                   const int32_t* hi_ptr, int hi_off) {
    if (kptr && *kptr == 0) return;
    // optional column window [lo, hi) read from device memory: tiles wholly
    // outside it would only add zero products
    if (lo_ptr && ((int64_t)blockIdx.x + 1) * 64 <= *lo_ptr) return;
    if (hi_ptr && (int64_t)blockIdx.x * 64 >= *hi_ptr + hi_off) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t m0 = (int64_t)blockIdx.y * 64 + (wid >> 1) * 32;
    const int64_t n0 = (int64_t)blockIdx.x * 64 + (wid & 1) * 32;
    const int r16 = lane & 15, g = lane >> 4;
    v4i acc[7][2][2];
#pragma unroll
    for (int s = 0; s < 7; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[s][i][j] = v4i{0, 0, 0, 0};
    const v4i zero = {0, 0, 0, 0};
    for (int kc = 0; kc < Kpad; kc += 64) {
        v4i a[4][2], b[4][2];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            const int64_t row = m0 + mi * 16 + r16;
#pragma unroll
            for (int l = 0; l < 4; ++l)
                a[l][mi] = row < Mrows ? *reinterpret_cast<const v4i*>(L + l * planeL + row * Kpad + kc + 16 * g) : zero;
        }
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
            const int64_t col = n0 + ni * 16 + r16;
#pragma unroll
            for (int l = 0; l < 4; ++l)
                b[l][ni] = col < Ncols ? *reinterpret_cast<const v4i*>(B + l * planeB + col * Kpad + kc + 16 * g) : zero;
        }
#pragma unroll
        for (int l1 = 0; l1 < 4; ++l1)
#pragma unroll
            for (int l2 = 0; l2 < 4; ++l2)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int ni = 0; ni < 2; ++ni)
                        acc[l1 + l2][mi][ni] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(a[l1][mi], b[l2][ni], acc[l1 + l2][mi][ni], 0, 0, 0);
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t row = m0 + mi * 16 + 4 * g + i;
                const int64_t col = n0 + ni * 16 + r16;
                if (row < Mrows && col < Ncols) {
                    int32_t s[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) s[q] = acc[q][mi][ni][i];
                    uint32_t* cp = C + row * ldc + col;
                    *cp = fold_sub(CZERO ? 0u : *cp, s, pd, pinv);
                }
            }
}

// ---------------------------------------------------------------------------
// the modular GEMM, second form: 512 threads, 128 x 64 tile per workgroup
// (8 waves of 32 x 32), operands staged through LDS in 64-deep K stages with
// double buffering, and a tile order that keeps each XCD's L2 on a small
// working set: consecutive workgroup ids go to different XCDs, so the tile
// list is first cut into 8 contiguous ranges (one per XCD), and each range is
// walked in groups of 8 tile-rows so that a B tile is reused by 8
// consecutive tiles while the 8 A tiles stay resident.
// ---------------------------------------------------------------------------

constexpr int GB_M = 128, GB_N = 64, GB_K = 64, GB_T = 512;
constexpr int GB_ROW = GB_K + 16;  // LDS bytes per staged row (padded against bank conflicts)
constexpr int GB_GROUP = 8;
constexpr int NUM_XCD = 8;

__device__ __forceinline__ bool gemm_tile_of(int64_t mt, int64_t nt, int64_t* tm, int64_t* tn) {
    const int64_t total = mt * nt;
    const int64_t per_xcd = (total + NUM_XCD - 1) / NUM_XCD;
    const int64_t bid = blockIdx.x;
    const int64_t t = (bid % NUM_XCD) * per_xcd + bid / NUM_XCD;
    if (t >= total) return false;
    const int64_t gsz = (int64_t)GB_GROUP * nt;
    const int64_t g = t / gsz, tt = t - g * gsz;
    const int64_t gm = min((int64_t)GB_GROUP, mt - g * GB_GROUP);
    *tm = g * GB_GROUP + tt % gm;
    *tn = tt / gm;
    return true;
}

inline unsigned gemm2_grid(int64_t Mrows, int64_t Ncols) {
    const int64_t total = ceil_div(Mrows, GB_M) * ceil_div(Ncols, GB_N);
    return (unsigned)(ceil_div(total, NUM_XCD) * NUM_XCD);
}

template <bool CZERO>
__global__ __launch_bounds__(GB_T) void k_gemm_modp2(uint32_t* __restrict__ C, int64_t ldc, int64_t Mrows,
                                                     int64_t Ncols, const int8_t* __restrict__ L, int64_t planeL,
                                                     const int8_t* __restrict__ B, int64_t planeB, int Kpad,
                                                     double pd, double pinv, const int32_t* kptr,
                                                     const int32_t* lo_ptr, const int32_t* hi_ptr, int hi_off) {
    __shared__ __attribute__((aligned(16))) int8_t sA[2][4][GB_M * GB_ROW];
    __shared__ __attribute__((aligned(16))) int8_t sB[2][4][GB_N * GB_ROW];
    if (kptr && *kptr == 0) return;
    int64_t tm, tn;
    if (!gemm_tile_of(ceil_div(Mrows, GB_M), ceil_div(Ncols, GB_N), &tm, &tn)) return;
    const int64_t m_base = tm * GB_M, n_base = tn * GB_N;
    if (lo_ptr && n_base + GB_N <= *lo_ptr) return;
    if (hi_ptr && n_base >= *hi_ptr + hi_off) return;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int r16 = lane & 15, g = lane >> 4;
    const v4i zero = {0, 0, 0, 0};

    // stage copy: A = 4 limbs x 128 rows x 4 sixteen-byte quarters (4 per thread),
    //             B = 4 limbs x 64 cols x 4 quarters (2 per thread)
    v4i ra[4], rb[2];
    auto gload = [&](int kc) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = tid + q * GB_T;
            const int l = c >> 9, rem = c & 511, row = rem >> 2, qu = rem & 3;
            const int64_t grow = m_base + row;
            ra[q] = grow < Mrows ? *reinterpret_cast<const v4i*>(L + l * planeL + grow * Kpad + kc + qu * 16) : zero;
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int c = tid + q * GB_T;
            const int l = c >> 8, rem = c & 255, col = rem >> 2, qu = rem & 3;
            const int64_t gcol = n_base + col;
            rb[q] = gcol < Ncols ? *reinterpret_cast<const v4i*>(B + l * planeB + gcol * Kpad + kc + qu * 16) : zero;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = tid + q * GB_T;
            const int l = c >> 9, rem = c & 511, row = rem >> 2, qu = rem & 3;
            *reinterpret_cast<v4i*>(&sA[buf][l][row * GB_ROW + qu * 16]) = ra[q];
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int c = tid + q * GB_T;
            const int l = c >> 8, rem = c & 255, col = rem >> 2, qu = rem & 3;
            *reinterpret_cast<v4i*>(&sB[buf][l][col * GB_ROW + qu * 16]) = rb[q];
        }
    };

    v4i acc[7][2][2];
#pragma unroll
    for (int s = 0; s < 7; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[s][i][j] = zero;

    const int nk = Kpad / GB_K;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kk = 0; kk < nk; ++kk) {
        const int buf = kk & 1;
        if (kk + 1 < nk) gload((kk + 1) * GB_K);
        v4i bf[4][2];
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
                bf[l][ni] = *reinterpret_cast<const v4i*>(&sB[buf][l][(wn * 32 + ni * 16 + r16) * GB_ROW + 16 * g]);
#pragma unroll
        for (int l1 = 0; l1 < 4; ++l1) {
            v4i af[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
                af[mi] = *reinterpret_cast<const v4i*>(&sA[buf][l1][(wm * 32 + mi * 16 + r16) * GB_ROW + 16 * g]);
#pragma unroll
            for (int l2 = 0; l2 < 4; ++l2)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int ni = 0; ni < 2; ++ni)
                        acc[l1 + l2][mi][ni] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(af[mi], bf[l2][ni], acc[l1 + l2][mi][ni], 0, 0, 0);
        }
        if (kk + 1 < nk) lstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t row = m_base + wm * 32 + mi * 16 + 4 * g + i;
                const int64_t col = n_base + wn * 32 + ni * 16 + r16;
                if (row < Mrows && col < Ncols) {
                    int32_t s[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) s[q] = acc[q][mi][ni][i];
                    uint32_t* cp = C + row * ldc + col;
                    *cp = fold_sub(CZERO ? 0u : *cp, s, pd, pinv);
                }
            }
}

// ---------------------------------------------------------------------------
// third form (used): 256 threads, 64 x 64 tile (4 waves of 32 x 32), 64 KiB of
// double-buffered LDS so that two independent workgroups share a CU and one's
// prologue / epilogue overlaps the other's MFMA work; the C tile is fetched
// before the K loop so its latency hides behind the MFMAs.  Same XCD-aware
// grouped tile order as the second form.
// ---------------------------------------------------------------------------

constexpr int G3_M = 64, G3_N = 64, G3_T = 256;

inline unsigned gemm3_grid(int64_t Mrows, int64_t Ncols) {
    const int64_t total = ceil_div(Mrows, G3_M) * ceil_div(Ncols, G3_N);
    return (unsigned)(ceil_div(total, NUM_XCD) * NUM_XCD);
}

template <bool CZERO>
__global__ __launch_bounds__(G3_T, 2) void k_gemm_modp3(uint32_t* __restrict__ C, int64_t ldc, int64_t Mrows,
                                                        int64_t Ncols, const int8_t* __restrict__ L,
                                                        int64_t planeL, const int8_t* __restrict__ B,
                                                        int64_t planeB, int Kpad, double pd, double pinv,
                                                        const int32_t* kptr, const int32_t* lo_ptr,
                                                        const int32_t* hi_ptr, int hi_off) {
    __shared__ __attribute__((aligned(16))) int8_t sA[2][4][G3_M * GB_ROW];
    __shared__ __attribute__((aligned(16))) int8_t sB[2][4][G3_N * GB_ROW];
    if (kptr && *kptr == 0) return;
    int64_t tm, tn;
    if (!gemm_tile_of(ceil_div(Mrows, G3_M), ceil_div(Ncols, G3_N), &tm, &tn)) return;
    const int64_t m_base = tm * G3_M, n_base = tn * G3_N;
    if (lo_ptr && n_base + G3_N <= *lo_ptr) return;
    if (hi_ptr && n_base >= *hi_ptr + hi_off) return;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int r16 = lane & 15, g = lane >> 4;
    const v4i zero = {0, 0, 0, 0};

    // C tile values of this lane: fetched while the last K stage runs on the
    // matrix cores (the stage registers are free by then)
    uint32_t cpre[2][2][4];
    auto cload = [&]() {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int64_t row = m_base + wm * 32 + mi * 16 + 4 * g + i;
                    const int64_t col = n_base + wn * 32 + ni * 16 + r16;
                    cpre[mi][ni][i] = (!CZERO && row < Mrows && col < Ncols) ? C[row * ldc + col] : 0u;
                }
    };

    // stage copy: A and B = 4 limbs x 64 rows x 4 sixteen-byte quarters (4 each per thread)
    v4i ra[4], rb[4];
    auto gload = [&](int kc) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = tid + q * G3_T;
            const int l = c >> 8, rem = c & 255, row = rem >> 2, qu = rem & 3;
            const int64_t grow = m_base + row, gcol = n_base + row;
            ra[q] = grow < Mrows ? *reinterpret_cast<const v4i*>(L + l * planeL + grow * Kpad + kc + qu * 16) : zero;
            rb[q] = gcol < Ncols ? *reinterpret_cast<const v4i*>(B + l * planeB + gcol * Kpad + kc + qu * 16) : zero;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = tid + q * G3_T;
            const int l = c >> 8, rem = c & 255, row = rem >> 2, qu = rem & 3;
            *reinterpret_cast<v4i*>(&sA[buf][l][row * GB_ROW + qu * 16]) = ra[q];
            *reinterpret_cast<v4i*>(&sB[buf][l][row * GB_ROW + qu * 16]) = rb[q];
        }
    };

    v4i acc[7][2][2];
#pragma unroll
    for (int s = 0; s < 7; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[s][i][j] = zero;

    const int nk = Kpad / GB_K;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kk = 0; kk < nk; ++kk) {
        const int buf = kk & 1;
        if (kk + 1 < nk) gload((kk + 1) * GB_K);
        else cload();
        v4i bf[4][2];
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
                bf[l][ni] = *reinterpret_cast<const v4i*>(&sB[buf][l][(wn * 32 + ni * 16 + r16) * GB_ROW + 16 * g]);
#pragma unroll
        for (int l1 = 0; l1 < 4; ++l1) {
            v4i af[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
                af[mi] = *reinterpret_cast<const v4i*>(&sA[buf][l1][(wm * 32 + mi * 16 + r16) * GB_ROW + 16 * g]);
#pragma unroll
            for (int l2 = 0; l2 < 4; ++l2)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int ni = 0; ni < 2; ++ni)
                        acc[l1 + l2][mi][ni] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(af[mi], bf[l2][ni], acc[l1 + l2][mi][ni], 0, 0, 0);
        }
        if (kk + 1 < nk) lstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t row = m_base + wm * 32 + mi * 16 + 4 * g + i;
                const int64_t col = n_base + wn * 32 + ni * 16 + r16;
                if (row < Mrows && col < Ncols) {
                    int32_t s[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) s[q] = acc[q][mi][ni][i];
                    C[row * ldc + col] = fold_sub(cpre[mi][ni][i], s, pd, pinv);
                }
            }
}

// ---------------------------------------------------------------------------
// inner-step GEMM on the panel buffer:  X[r, :] -= X[r, c0 + Cp] U  for rows
// that are not pivots of this step, X[r, :] = U[slot(r), :] for those that
// are.  K = 64 (k <= 32 pivots, zero padded), one LDS stage.  The A operand's
// limbs are built from X on the fly (no separate plane buffer), B is the limb
// planes of U written by k_make_u.  Tiles outside the live column window
// [c0, BO + q0) only meet zeros and are skipped.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(G3_T, 2) void k_gemm_inner(uint32_t* __restrict__ X, int64_t N,
                                                        const int8_t* __restrict__ Bpl, int64_t planeB,
                                                        const uint32_t* __restrict__ U,
                                                        const BlockState* __restrict__ bs,
                                                        const int32_t* __restrict__ epoch_of, int epoch,
                                                        const int32_t* __restrict__ slot_of, uint32_t p, double pd,
                                                        double pinv) {
    __shared__ __attribute__((aligned(16))) int8_t sA[4][G3_M * GB_ROW];
    __shared__ __attribute__((aligned(16))) int8_t sB[4][G3_N * GB_ROW];
    __shared__ int s_cp[CI];
    const int k = bs->k;
    if (k == 0) return;
    int64_t tm, tn;
    if (!gemm_tile_of(ceil_div(N, G3_M), XLD / G3_N, &tm, &tn)) return;
    const int64_t m_base = tm * G3_M, n_base = tn * G3_N;
    const int c0 = bs->c0_cur;
    if (n_base + G3_N <= c0 || n_base >= BO + bs->q0) return;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int r16 = lane & 15, g = lane >> 4;
    if (tid < k) s_cp[tid] = bs->Cp[tid];
    __syncthreads();
    // A: thread -> (row = tid / 4, 16 consecutive k), values X[row, c0 + Cp[t]]
    {
        const int rl = tid >> 2, kq = (tid & 3) * 16;
        const int64_t row = m_base + rl;
        int8_t dg[4][16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int t = kq + i;
            int8_t d[4] = {0, 0, 0, 0};
            if (t < k && row < N) balanced_digits(centred(X[row * XLD + c0 + s_cp[t]], p), d);
#pragma unroll
            for (int l = 0; l < 4; ++l) dg[l][i] = d[l];
        }
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            v4i w;
            __builtin_memcpy(&w, dg[l], 16);
            *reinterpret_cast<v4i*>(&sA[l][rl * GB_ROW + kq]) = w;
        }
    }
    // B: copy the planes of U for this tile's 64 columns
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = tid + q * G3_T;
        const int l = c >> 8, rem = c & 255, col = rem >> 2, qu = rem & 3;
        *reinterpret_cast<v4i*>(&sB[l][col * GB_ROW + qu * 16]) =
            *reinterpret_cast<const v4i*>(Bpl + l * planeB + (n_base + col) * KPI + qu * 16);
    }
    __syncthreads();
    const v4i zero = {0, 0, 0, 0};
    v4i acc[7][2][2];
#pragma unroll
    for (int s = 0; s < 7; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[s][i][j] = zero;
    v4i bf[4][2];
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
            bf[l][ni] = *reinterpret_cast<const v4i*>(&sB[l][(wn * 32 + ni * 16 + r16) * GB_ROW + 16 * g]);
#pragma unroll
    for (int l1 = 0; l1 < 4; ++l1) {
        v4i af[2];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
            af[mi] = *reinterpret_cast<const v4i*>(&sA[l1][(wm * 32 + mi * 16 + r16) * GB_ROW + 16 * g]);
#pragma unroll
        for (int l2 = 0; l2 < 4; ++l2)
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni)
                    acc[l1 + l2][mi][ni] =
                        __builtin_amdgcn_mfma_i32_16x16x64_i8(af[mi], bf[l2][ni], acc[l1 + l2][mi][ni], 0, 0, 0);
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t row = m_base + wm * 32 + mi * 16 + 4 * g + i;
            if (row >= N) continue;
            const bool pivot_row = epoch_of[row] == epoch;
            const int slot = pivot_row ? slot_of[row] : 0;
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const int64_t col = n_base + wn * 32 + ni * 16 + r16;
                uint32_t* cp = X + row * XLD + col;
                if (pivot_row) {
                    *cp = U[slot * XLD + col];
                } else {
                    int32_t sv[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) sv[q] = acc[q][mi][ni][i];
                    *cp = fold_sub(*cp, sv, pd, pinv);
                }
            }
        }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

struct BlockedBuffers {
    uint32_t* X;
    int32_t* slot_of;   // pivot slot of a row within its inner step
    int32_t* epoch_of;
    BlockState* bs;
    int8_t* planesL;  // [4][N][256]
    int64_t strideL;
    int8_t* planesY;  // [4][256][256]
    int64_t strideY;
    int8_t* planesB;  // [4][max(M, XLD)][256]
    int64_t strideB;
    uint32_t* U;      // [CI][XLD]
    uint32_t* Gq;     // [256][ldg]
    int64_t ldg;
};

static int blocked_alloc(DeviceContext* ctx, int64_t N, int64_t M, BlockedBuffers* b) {
    void* p = nullptr;
    LINAC_TRY(ctx->get(SLOT_PANEL_A, (size_t)N * XLD * 4, &p));
    b->X = (uint32_t*)p;
    b->strideL = round_up(N * 256, 256);
    b->strideY = 256 * 256;
    LINAC_TRY(ctx->get(SLOT_PANEL_B, (size_t)(4 * b->strideL + 4 * b->strideY), &p));
    b->planesL = (int8_t*)p;
    b->planesY = b->planesL + 4 * b->strideL;
    const int64_t nb = std::max<int64_t>(M, XLD);
    b->strideB = round_up(nb * 256, 256);
    b->ldg = round_up(M, 64);
    const size_t bytesB = (size_t)(4 * b->strideB);
    const size_t bytesU = (size_t)CI * XLD * 4;
    const size_t bytesG = (size_t)256 * b->ldg * 4;
    LINAC_TRY(ctx->get(SLOT_PANEL_C, bytesB + bytesU + bytesG, &p));
    b->planesB = (int8_t*)p;
    b->U = (uint32_t*)((char*)p + bytesB);
    b->Gq = (uint32_t*)((char*)p + bytesB + bytesU);
    const size_t aux = 2 * (size_t)round_up(4 * N, 256) + sizeof(BlockState) + 256;
    LINAC_TRY(ctx->get(SLOT_AUX, aux, &p));
    b->slot_of = (int32_t*)p;
    b->epoch_of = (int32_t*)((char*)p + round_up(4 * N, 256));
    b->bs = (BlockState*)((char*)p + 2 * round_up(4 * N, 256));
    return LINAC_OK;
}

static unsigned grid1(int64_t n, int64_t per = 256, int64_t cap = 65536) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cap, ceil_div(n, per)));
}

// One inner-step group.  Which inner panel it works on is decided on the
// device (BlockState::c0_next); a group launched after the outer panel is
// complete does nothing.
static int inner_group(const BlockedBuffers& b, int64_t N, int epoch, int64_t J, const FieldU32& f, double pd,
                       double pinv, hipStream_t s, const ElimWorkspace& ws) {
    k_inner_factor<<<1, IF_T, 0, s>>>(b.X, N, ws.used, b.slot_of, b.epoch_of, epoch, J, f, ws.st, ws.piv_rows,
                                     ws.piv_cols, b.bs);
    k_make_u<<<XLD / 4, dim3(64, 4), 0, s>>>(b.X, f, b.bs, b.U, b.planesB, b.strideB);
    k_gemm_inner<<<(unsigned)(ceil_div(ceil_div(N, G3_M) * (XLD / G3_N), NUM_XCD) * NUM_XCD), G3_T, 0, s>>>(
        b.X, N, b.planesB, b.strideB, b.U, b.bs, b.epoch_of, epoch, b.slot_of, f.p, pd, pinv);
    LINAC_HIP(hipGetLastError());
    return LINAC_OK;
}

static int blocked_run_u32(uint32_t* A, int64_t ld, int64_t N, int64_t M, const FieldU32& f, hipStream_t s,
                           const ElimWorkspace& ws, DeviceContext* ctx, Profiler* prof) {
    BlockedBuffers b;
    LINAC_TRY(blocked_alloc(ctx, N, M, &b));
    const double pd = (double)f.p, pinv = 1.0 / (double)f.p;
    k_epoch_init<<<grid1(N), 256, 0, s>>>(b.epoch_of, N);
    int epoch = 0;
    for (int64_t J = 0; J < M; J += BO) {
        const int Bcur = (int)std::min<int64_t>(BO, M - J);
        k_panel_load<<<grid1(N * XLD), 256, 0, s>>>(b.X, A, ld, N, J, Bcur);
        k_block_state_init<<<1, 1, 0, s>>>(b.bs, Bcur, J == 0);
        const int e0 = epoch;
        const int ninner = (int)ceil_div(Bcur, BI);
        for (int i = 0; i < ninner; ++i) LINAC_TRY(inner_group(b, N, epoch++, J, f, pd, pinv, s, ws));
        // one synchronisation per outer panel; extra groups only when an inner
        // panel needed more than one round (rank-deficient panels)
        BlockState hs;
        for (int64_t guard = 0;; ++guard) {
            LINAC_HIP(hipMemcpyAsync(&hs, b.bs, offsetof(BlockState, R), hipMemcpyDeviceToHost, s));
            LINAC_HIP(hipStreamSynchronize(s));
            if (hs.c0_next >= Bcur) break;
            if (guard > (int64_t)BO * (BI + 1)) return fail(LINAC_ERR_DEVICE, "blocked elimination did not converge");
            const int extra = std::max(1, (int)ceil_div(Bcur - hs.c0_next, BI));
            for (int i = 0; i < extra; ++i) LINAC_TRY(inner_group(b, N, epoch++, J, f, pd, pinv, s, ws));
        }
        LINAC_HIP(hipMemcpyAsync(ctx->pinned, &ws.st->rank, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
        const int64_t rank_end = ctx->pinned[0];
        const int q0 = hs.q0;
        const int64_t rank_start = rank_end - q0;
        const int64_t T0 = J + Bcur, nT = M - T0;
        if (q0 > 0 && nT > 0) {
            const int Kq = (int)round_up(q0, 64);
            const int64_t* Q = ws.piv_rows + rank_start;
            k_make_lhs<<<grid1(N * Kq, 256, 1 << 30), 256, 0, s>>>(b.X, XLD, N, nullptr, BO, nullptr, nullptr, nullptr, q0, Kq,
                                                                    b.epoch_of, e0, epoch, 1, f.p, b.planesL,
                                                                    b.strideL);
            k_make_lhs<<<grid1((int64_t)q0 * Kq, 256, 1 << 30), 256, 0, s>>>(b.X, XLD, q0, Q, BO, nullptr, nullptr,
                                                                              nullptr, q0, Kq, nullptr, 0, 0, 1, f.p,
                                                                              b.planesY, b.strideY);
            k_make_rhs<<<dim3((unsigned)ceil_div(nT, 64), (unsigned)(Kq / 64)), 256, 0, s>>>(
                A, ld, Q, q0, T0, nT, Kq, f.p, b.planesB, b.strideB);
            k_gemm_modp3<true><<<gemm3_grid(q0, nT), G3_T, 0, s>>>(b.Gq, b.ldg, q0, nT, b.planesY, b.strideY,
                                                                   b.planesB, b.strideB, Kq, pd, pinv, nullptr,
                                                                   nullptr, nullptr, 0);
            if (prof)
                LINAC_TRY(prof->sample_begin(s, J, 8.0 * (double)N * (double)nT,
                                             (double)(N - q0) * (double)q0 * (double)nT));
            k_gemm_modp3<false><<<gemm3_grid(N, nT), G3_T, 0, s>>>(A + T0, ld, N, nT, b.planesL, b.strideL,
                                                                   b.planesB, b.strideB, Kq, pd, pinv, nullptr,
                                                                   nullptr, nullptr, 0);
            if (prof) LINAC_TRY(prof->sample_end(s));
            k_scatter_rows<<<dim3(grid1(nT, 256, 64), (unsigned)q0), 256, 0, s>>>(A, ld, Q, nullptr, q0, T0, nT,
                                                                                b.Gq, b.ldg);
        }
        k_panel_store<<<grid1(N * Bcur), 256, 0, s>>>(A, b.X, ld, N, J, Bcur);
        LINAC_HIP(hipGetLastError());
    }
    LINAC_HIP(hipStreamSynchronize(s));
    if (getenv("LINAC_DEBUG_TIMING")) {
        unsigned long long tp[8];
        LINAC_HIP(hipMemcpy(tp, b.bs->tphase, sizeof(tp), hipMemcpyDeviceToHost));
        int rate_khz = 0;
        LINAC_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
        const double us = 1e3 / (double)rate_khz;
        fprintf(stderr, "k_inner_factor phases over %llu calls (us/call): scan %.2f load %.2f gj %.2f norm %.2f out %.2f\n",
                tp[5], tp[0] * us / tp[5], tp[1] * us / tp[5], tp[2] * us / tp[5], tp[3] * us / tp[5], tp[4] * us / tp[5]);
    }
    return LINAC_OK;
}

template <class F>
int blocked_dispatch(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                     const ElimWorkspace& ws, DeviceContext* ctx, bool forced, Profiler* prof) {
    if constexpr (std::is_same<F, FieldU32>::value) {
        if (f.p > LIMB_P_MAX) return NOT_APPLICABLE;
        if (!forced && (N < 1024 || M < 512)) return NOT_APPLICABLE;
        return blocked_run_u32(A, ld, N, M, f, s, ws, ctx, prof);
    } else {
        (void)A; (void)ld; (void)N; (void)M; (void)f; (void)s; (void)ws; (void)ctx; (void)forced; (void)prof;
        return NOT_APPLICABLE;
    }
}

template int blocked_dispatch<FieldU32>(uint32_t*, int64_t, int64_t, int64_t, const FieldU32&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*);
template int blocked_dispatch<FieldU64>(uint64_t*, int64_t, int64_t, int64_t, const FieldU64&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*);
template int blocked_dispatch<FieldGF2>(uint32_t*, int64_t, int64_t, int64_t, const FieldGF2&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*);

}  // namespace linac
