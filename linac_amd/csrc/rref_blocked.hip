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

__global__ void k_block_state_init(BlockState* bs, int bcur) {
    bs->k = 0;
    bs->q0 = 0;
    bs->c0_next = 0;
    bs->c0_cur = 0;
    bs->bcur = bcur;
}

// flags[r] = row r is unused and has a non-zero entry in the next inner panel
__global__ __launch_bounds__(256) void k_row_flags(const uint32_t* __restrict__ X, int64_t N,
                                                   const BlockState* __restrict__ bs,
                                                   const uint8_t* __restrict__ used, uint8_t* __restrict__ flags) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = bs->c0_next;
    if (r >= N || c0 >= bs->bcur) return;
    const int w = min(BI, bs->bcur - c0);
    uint8_t f = 0;
    if (!used[r]) {
        const uint32_t* row = X + r * XLD + c0;
        uint32_t acc = 0;
        for (int j = 0; j < w; ++j) acc |= row[j];
        f = acc != 0u;
    }
    flags[r] = f;
}

// ---------------------------------------------------------------------------
// inner factorisation: one workgroup
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_inner_factor(uint32_t* __restrict__ X, int64_t N,
                                                      const uint8_t* __restrict__ flags, uint8_t* __restrict__ used,
                                                      int32_t* __restrict__ epoch_of, int epoch,
                                                      int64_t colbase, FieldU32 f, ElimState* st,
                                                      int64_t* __restrict__ piv_rows, int64_t* __restrict__ piv_cols,
                                                      BlockState* bs) {
    constexpr int W2 = BI + CI;
    __shared__ uint32_t aug[CI][W2 + 1];
    __shared__ int64_t cand[CI + 1];
    __shared__ int orig[CI];
    __shared__ int pcol[CI];
    __shared__ uint32_t am[CI];
    __shared__ uint32_t invm[CI];
    __shared__ int s_total, s_piv;
    __shared__ int s_wcnt[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
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
    for (int64_t base = st->first_unused; base < N; base += 256) {
        const int64_t r = base + tid;
        const bool fl = r < N && flags[r];
        const unsigned long long m = __ballot(fl);
        if (lane == 0) s_wcnt[wid] = __popcll(m);
        __syncthreads();
        int before = s_total;
        for (int q = 0; q < wid; ++q) before += s_wcnt[q];
        const int pos = before + __popcll(m & ((1ull << lane) - 1ull));
        if (fl && pos <= CI) cand[pos] = r;
        __syncthreads();
        if (tid == 0) s_total += s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
        __syncthreads();
        if (s_total > CI) break;
    }
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

    // 2. [X[S, c0:c0+w] | I] in LDS
    for (int x = tid; x < nC * W2; x += 256) {
        const int i = x / W2, j = x - i * W2;
        uint32_t v;
        if (j < w) v = X[cand[i] * XLD + c0 + j];
        else if (j < BI) v = 0u;
        else v = (j - BI == i) ? 1u : 0u;
        aug[i][j] = v;
    }
    if (tid < nC) orig[tid] = tid;
    __syncthreads();

    // 3. fraction-free Gauss-Jordan: row_i <- h * row_i - a_i * row_t.
    //    Rows stay scalar multiples of the exact Gauss-Jordan rows; pivot rows
    //    are normalised once at the end, so no inverse sits on the chain.
    int t = 0;
    for (int c = 0; c < w && t < nC; ++c) {
        if (wid == 0) {
            const bool nz = lane >= t && lane < nC && aug[lane < CI ? lane : 0][c] != 0u;
            const unsigned long long m = __ballot(nz);
            if (lane == 0) s_piv = m ? (__ffsll((long long)m) - 1) : -1;
        }
        __syncthreads();
        const int pv = s_piv;
        if (pv < 0) {
            __syncthreads();
            continue;
        }
        if (pv != t) {
            for (int j = tid; j < W2; j += 256) {
                const uint32_t a = aug[t][j];
                aug[t][j] = aug[pv][j];
                aug[pv][j] = a;
            }
            if (tid == 0) {
                const int o = orig[t];
                orig[t] = orig[pv];
                orig[pv] = o;
            }
        }
        __syncthreads();
        if (tid < nC) am[tid] = f.to_mont(aug[tid][c]);
        __syncthreads();
        const uint32_t hm = am[t];
        for (int x = tid; x < nC * W2; x += 256) {
            const int i = x / W2, j = x - i * W2;
            if (i != t) aug[i][j] = f.sub(f.mont(hm, aug[i][j]), f.mont(am[i], aug[t][j]));
        }
        if (tid == 0) pcol[t] = c;
        ++t;
        __syncthreads();
    }
    const int k = t;

    // 4. normalise the k pivot rows (k independent inverses in parallel)
    if (tid < k) invm[tid] = f.to_mont(f.inverse(aug[tid][pcol[tid]]));
    __syncthreads();
    for (int x = tid; x < k * W2; x += 256) {
        const int i = x / W2, j = x - i * W2;
        aug[i][j] = f.mont(invm[i], aug[i][j]);
    }
    __syncthreads();

    // 5. outputs: Apiv^{-1} (Montgomery form), pivot rows/columns, bookkeeping
    const int64_t rank0 = st->rank;
    const int q0 = bs->q0;
    for (int x = tid; x < k * k; x += 256) {
        const int a = x / k, b = x - a * k;
        bs->apinv[a * CI + b] = f.to_mont(aug[a][BI + orig[b]]);
    }
    if (tid < k) {
        const int64_t r = cand[orig[tid]];
        bs->R[tid] = r;
        bs->Cp[tid] = pcol[tid];
        used[r] = 1;
        epoch_of[r] = epoch;
        piv_rows[rank0 + tid] = r;
        piv_cols[rank0 + tid] = colbase + c0 + pcol[tid];
        X[r * XLD + BO + q0 + tid] = 1u;  // Z[r, slot] = 1
    }
    __syncthreads();
    if (tid == 0) {
        st->rank = rank0 + k;
        bs->k = k;
        bs->q0 = q0 + k;
        bs->c0_cur = c0;
        bs->c0_next = (more && k < w) ? c0 : c0 + BI;
        int64_t q = st->first_unused;
        while (q < N && used[q]) ++q;
        st->first_unused = q;
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
// host driver
// ---------------------------------------------------------------------------

struct BlockedBuffers {
    uint32_t* X;
    uint8_t* flags;
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
    const size_t aux = (size_t)round_up(N, 256) + (size_t)round_up(4 * N, 256) + sizeof(BlockState) + 256;
    LINAC_TRY(ctx->get(SLOT_AUX, aux, &p));
    b->flags = (uint8_t*)p;
    b->epoch_of = (int32_t*)((char*)p + round_up(N, 256));
    b->bs = (BlockState*)((char*)p + round_up(N, 256) + round_up(4 * N, 256));
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
    k_row_flags<<<grid1(N, 256, 1 << 30), 256, 0, s>>>(b.X, N, b.bs, ws.used, b.flags);
    k_inner_factor<<<1, 256, 0, s>>>(b.X, N, b.flags, ws.used, b.epoch_of, epoch, J, f, ws.st, ws.piv_rows,
                                     ws.piv_cols, b.bs);
    k_make_u<<<XLD / 4, dim3(64, 4), 0, s>>>(b.X, f, b.bs, b.U, b.planesB, b.strideB);
    k_make_lhs<<<grid1(N * KPI, 256, 1 << 30), 256, 0, s>>>(b.X, XLD, N, nullptr, 0, &b.bs->c0_cur, b.bs->Cp,
                                                             &b.bs->k, 0, KPI, b.epoch_of, epoch, epoch + 1, 0, f.p,
                                                             b.planesL, b.strideL);
    dim3 gg(XLD / 64, (unsigned)ceil_div(N, 64));
    k_gemm_modp<false><<<gg, 256, 0, s>>>(b.X, XLD, N, XLD, b.planesL, b.strideL, b.planesB, b.strideB, KPI, f.p,
                                          pd, pinv, &b.bs->k, &b.bs->c0_cur, &b.bs->q0, BO);
    k_scatter_rows<<<dim3(2, CI), 256, 0, s>>>(b.X, XLD, b.bs->R, &b.bs->k, 0, 0, XLD, b.U, XLD);
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
        k_block_state_init<<<1, 1, 0, s>>>(b.bs, Bcur);
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
            k_gemm_modp<true><<<dim3((unsigned)ceil_div(nT, 64), (unsigned)ceil_div(q0, 64)), 256, 0, s>>>(
                b.Gq, b.ldg, q0, nT, b.planesY, b.strideY, b.planesB, b.strideB, Kq, f.p, pd, pinv, nullptr, nullptr,
                nullptr, 0);
            if (prof)
                LINAC_TRY(prof->sample_begin(s, J, 8.0 * (double)N * (double)nT,
                                             (double)(N - q0) * (double)q0 * (double)nT));
            k_gemm_modp<false><<<dim3((unsigned)ceil_div(nT, 64), (unsigned)ceil_div(N, 64)), 256, 0, s>>>(
                A + T0, ld, N, nT, b.planesL, b.strideL, b.planesB, b.strideB, Kq, f.p, pd, pinv, nullptr, nullptr,
                nullptr, 0);
            if (prof) LINAC_TRY(prof->sample_end(s));
            k_scatter_rows<<<dim3(grid1(nT, 256, 64), (unsigned)q0), 256, 0, s>>>(A, ld, Q, nullptr, q0, T0, nT,
                                                                                b.Gq, b.ldg);
        }
        k_panel_store<<<grid1(N * Bcur), 256, 0, s>>>(A, b.X, ld, N, J, Bcur);
        LINAC_HIP(hipGetLastError());
    }
    LINAC_HIP(hipStreamSynchronize(s));
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
