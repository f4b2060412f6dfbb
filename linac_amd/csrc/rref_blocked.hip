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
// Batches: every kernel takes an array of per-matrix descriptors (BMat) and
// serves a whole set of matrices of any shapes and primes in one launch; each
// matrix keeps its own panel position, so matrices can join or leave between
// steps (continuous batching).  A single solve is a batch of one.
//
// Rows are never swapped; `used` marks pivot rows, the pivot list records
// (row, column) and the shared finalisation gathers rows in column order.
// Pivot rows come out normalised.  Replaces the column loop of
// linac/pycuda_row_reduce.py:94-131 for large GF(p) matrices.
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <iterator>
#include <map>
#include <type_traits>
#include <vector>

#include "gf_common.hpp"
#include "modp_gemm.hpp"

namespace linac {

constexpr int BO = 256;        // outer panel width (GEMM depth)
constexpr int BI = 32;         // inner panel width
constexpr int CI = 32;         // candidate rows per inner step
constexpr int XLD = 2 * BO;    // leading dimension of the panel buffer [P | Z]
constexpr int KPI = 64;        // padded depth of the inner GEMM

// The two storage widths: element type, field, signed type of centred values
// and the number of balanced base-256 digits (int8 limb planes) per entry.
// 32-bit: p <= 4278124287, 4 digits, 16 limb products on 7 diagonals;
// 64-bit: 2^32 <= p < 2^63, 8 digits, 64 limb products on 15 diagonals.
// IF_T: threads of the inner factorisation; IF_WAVES: its minimum waves per
// SIMD, i.e. its register cap (512 / IF_WAVES).  The 64-bit one runs 8 waves
// held to 128 registers so that they fit in the space one GEMM workgroup frees
// on a CU (the trailing update of the previous panel runs beside it and
// otherwise refills every freed slot first); 16 waves would have to live in
// 64 registers, which spills.
struct LimbsU32 {
    using T = uint32_t;
    using F = FieldU32;
    using S = int32_t;
    static constexpr int NL = 4;
    static constexpr int IF_T = 1024;
    static constexpr int IF_WAVES = 1;
};
struct LimbsU64 {
    using T = uint64_t;
    using F = FieldU64;
    using S = int64_t;
    static constexpr int NL = 8;
    static constexpr int IF_T = 512;
    static constexpr int IF_WAVES = 4;
};

template <class T>
struct BlockState {
    int32_t k;           // pivots found by the last inner step
    int32_t q0;          // pivot slots used in the current outer panel
    int32_t c0_next;     // first column (in the panel) of the next inner step
    int32_t c0_cur;      // first column of the last inner step
    int32_t bcur;        // width of the current outer panel
    int32_t pad_;
    int64_t lo0;         // first unused row when the outer panel was loaded
    int64_t R[CI];       // pivot rows of the last inner step (pivot order)
    int32_t Cp[CI];      // their columns relative to the inner panel start
    T apinv[CI * CI];    // Montgomery form of (X[R, c0 + Cp])^{-1}
    unsigned long long tphase[8];  // accumulated wall-clock ticks per phase of k_inner_factor
};

// ---------------------------------------------------------------------------
// Per-matrix descriptor.  Every kernel below serves a whole batch: the matrix
// index comes from blockIdx.y (uniform grids) or from a prefix sum of tile
// counts (GEMMs).  The host rewrites the descriptors before each step.
// ---------------------------------------------------------------------------

template <class LT>
struct BMat {
    using T = typename LT::T;
    T* A;
    int64_t ld, N, M;
    typename LT::F f;
    int32_t pad0;
    double pd, pinv;
    T* X;               // [N][XLD] panel buffer [P | Z]
    int8_t* planesL;    // [NL][N][2 BO]  left operand of the trailing update, row stride Kq
    int8_t* planesW;    // [NL][BO][BO]   right operand of the composite transformation, row stride Kq - KqA
    int8_t* planesB;    // [NL][max(M, XLD)][2 BO] right operand of the trailing update, row stride Kq
    int8_t* planesAi;   // [NL][N][KPI]   left operand of the inner GEMM
    int8_t* planesUi;   // [NL][XLD][KPI] right operand of the inner GEMM
    T* U;               // [CI][XLD]
    T* W;               // [N][BO]  transformation of the deferred panel (ld BO)
    int8_t* planesL1;   // [NL][N][BO]  operands of a deferring panel's own update (the
    int8_t* planesB1;   // [NL][BO][BO]  composite operands stay live for its second half)
    int64_t strideL, strideB, strideAi, strideL1;
    BlockState<T>* bs;
    ElimState* st;
    int64_t* piv_rows;
    int64_t* piv_cols;
    uint8_t* used;
    int32_t* epoch_of;
    int32_t* slot_of;
    // step parameters
    int64_t J;          // first column of the current outer panel
    int32_t bcur;       // its width
    int32_t q0;         // pivots it produced (outer update)
    int32_t kA;         // pivots of the deferred previous panel (0: none)
    int32_t KqA;        // kA rounded up to 64: depth of the deferred part of the trailing update
    int32_t Kq;         // KqA + q0 rounded up to 64 (GEMM depth, row stride of planesL / planesB)
    int32_t save_w;     // this panel's update beyond the next panel is deferred: keep W
    int32_t e0;         // first epoch of this outer panel
    int32_t full;       // every row is a pivot row: update every trailing column now, then stop
    int64_t rank_start; // index of its first pivot in piv_rows (the deferred panel's kA precede it)
    int64_t T0, nT;     // trailing columns [T0, T0 + nT)
    int64_t r0;         // first row of the trailing update, the panel write-back and the left operand
                        // (a multiple of 64; every row above it is a pivot row of an earlier panel)
    int64_t Jst, bst;   // panel [Jst, Jst + bst) whose write-back the next panel load does (bst = 0: none)
    int64_t rank0;      // rank when the step began
    int32_t store_now;  // this step writes its panel back itself (the item's last step)
    int32_t pad1;
    int64_t* cand;      // [BO] rows of the compact candidate block (full-rank panel fast path)
};

constexpr int64_t STRIDE_W = (int64_t)BO * BO;

// Limb planes of the trailing GEMMs' operands (planesL / planesW / planesB /
// planesL1 / planesB1 and the planes of the echelon completion and of
// linac_gemm_sub) are stored in tiles of 16 rows x 64 bytes of depth: byte kb
// of row r of a plane with row length ld (a multiple of 64) sits at
//   (r / 16) * 16 ld + (kb / 64) * 1024 + (r % 16) * 64 + kb % 64,
// so the 16 rows x 64 bytes that one direct-to-LDS load instruction stages
// are one contiguous kilobyte (eight whole 128-byte lines) instead of 16
// pieces of 64 bytes ld apart, and a stage advances by 1024 bytes.  A depth
// offset kb0 (a multiple of 64) is a pointer offset of 16 kb0 bytes; a row
// offset r0 (a multiple of 16) one of r0 ld bytes.  Planes hold their row
// count rounded up to 16 (prows).
__host__ __device__ __forceinline__ int64_t ptile(int64_t r, int64_t kb, int64_t ld) {
    return (r >> 4) * (ld << 4) + ((kb >> 6) << 10) + ((r & 15) << 6) + (kb & 63);
}
__host__ __device__ __forceinline__ int64_t prows(int64_t rows) { return (rows + 15) & ~(int64_t)15; }

// LINAC_DEBUG_STEPS: one line per outer panel on stderr (row offsets, rank,
// host time of the step so far and its wait for the status read)
static bool debug_steps() {
    static const bool on = getenv("LINAC_DEBUG_STEPS") != nullptr;
    return on;
}

// ---------------------------------------------------------------------------
// panel buffer management
// ---------------------------------------------------------------------------

// Panel copies in 16-byte vectors (J is a multiple of the panel width, so a
// vector that ends within the panel stays within its row); at the right edge
// of A the last, partial vector of a row is read element by element.  Rows
// need not be 16-byte aligned (a batch works in place on the caller's
// matrix, whatever its row stride): global loads accept any dword-aligned
// address.
template <class LT>
__device__ static void block_state_init(const BMat<LT>& d, int reset_timers);

template <class LT>
__device__ static void cand_gather_body(const BMat<LT>& d, const BMat<LT>& c);

// (workgroup 0 of each matrix also resets its inner-step state; with Dc, the
// last workgroup gathers the full-rank fast path's candidate rows)
template <class LT>
__global__ __launch_bounds__(256) void k_panel_load(const BMat<LT>* __restrict__ D, int reset_timers,
                                                    const BMat<LT>* __restrict__ Dc = nullptr) {
    using T = typename LT::T;
    constexpr int PER = 16 / (int)sizeof(T);
    constexpr int VPR = XLD / PER;  // vectors per panel-buffer row
    const BMat<LT>& d = D[blockIdx.y];
    const unsigned nload = gridDim.x - (Dc ? 1u : 0u);
    if (Dc && blockIdx.x == nload) {
        cand_gather_body<LT>(d, Dc[blockIdx.y]);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        block_state_init<LT>(d, reset_timers);
        d.bs->lo0 = d.st->first_unused;  // no pivot of this panel lies above it
    }
    // rows from r0 on (the previous step's: rows above it are pivot rows
    // that no later update touches, so their panel-buffer rows are never read
    // back; the candidates, the fast path's GEMM rows and everything written
    // back lie at or beyond it)
    // The previous panel's write-back (bst > 0) goes first, element by element
    // from the same vector of the buffer: the previous step's rows from r0 on,
    // which are the rows loaded here.
    const int64_t total = (d.N - d.r0) * VPR;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)nload * blockDim.x) {
        const int64_t i = x / VPR, r = d.r0 + i;
        const int j = (int)(x - i * VPR) * PER;
        if (j < d.bst) {
            const v4i raw = gld<v4i>(d.X + r * XLD + j);
            Vec16<T> o;
            __builtin_memcpy(&o, &raw, 16);
#pragma unroll
            for (int e = 0; e < PER; ++e)
                if (j + e < d.bst) d.A[r * d.ld + d.Jst + j + e] = o.v[e];
        }
        Vec16<T> v = {};
        if (j + PER <= d.bcur) {
            const v4i raw = gld<v4i>(d.A + r * d.ld + d.J + j);
            __builtin_memcpy(&v, &raw, 16);
        } else if (j < d.bcur) {
            // the vector would run past the panel's last column (at the right
            // edge of A, past the end of a row, and of the matrix on its last row)
#pragma unroll
            for (int e = 0; e < PER; ++e)
                if (j + e < d.bcur) v.v[e] = d.A[r * d.ld + d.J + j + e];
        }
        v4i out;
        __builtin_memcpy(&out, &v, 16);
        gst<v4i>(d.X + r * XLD + j, out);
    }
}

template <class LT>
__global__ __launch_bounds__(256) void k_panel_store(const BMat<LT>* __restrict__ D) {
    // one panel row per pass of the workgroup, one element per thread (this
    // measured faster than 16-byte vectors: 12 against 20 us at N = 20000)
    static_assert(BO == 256, "one thread per panel column");
    const BMat<LT>& d = D[blockIdx.y];
    const int j = threadIdx.x;
    if (j >= d.bcur || !d.store_now) return;  // (otherwise the next panel load writes it back)
    // (rows above r0 keep the values they had: the trailing update skips them too)
    for (int64_t r = d.r0 + blockIdx.x; r < d.N; r += gridDim.x) d.A[r * d.ld + d.J + j] = d.X[r * XLD + j];
}

__global__ __launch_bounds__(256) void k_epoch_init(int32_t* epoch, int64_t N) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += (int64_t)gridDim.x * blockDim.x)
        epoch[r] = -1;
}

template <class LT>
__device__ static void block_state_init(const BMat<LT>& d, int reset_timers) {
    auto* bs = d.bs;
    if (reset_timers)
        for (int i = 0; i < 8; ++i) bs->tphase[i] = 0;
    bs->k = 0;
    bs->q0 = 0;
    bs->c0_next = 0;
    bs->c0_cur = 0;
    bs->bcur = d.bcur;
}

// ---------------------------------------------------------------------------
// inner factorisation: one workgroup per matrix
// ---------------------------------------------------------------------------

static_assert(BI + CI == 64, "one augmented column per lane");
static_assert((CI & (CI - 1)) == 0 && CI <= 32, "pivot slots index with a mask; the inner GEMM holds k <= 32 in its first two lane groups");

// Candidate rows live in LDS as [panel columns | identity] (32-bit storage;
// 64-bit storage uses a compact form of half the width, described at its
// elimination loop); wave w owns rows w and w + 16 (w, w + 8, w + 16, w + 24
// with the eight waves of 64-bit storage).  (Each column step is a chain of dependent LDS reads and
// products: four waves with eight rows each took twice as long as sixteen
// with two, which hide each other's latency.)  Per column every wave ballots
// the column itself (all waves see the same data, so they agree on the pivot
// without a barrier), then
// updates its own rows with lane j handling column j.  Within a wave the
// reads of a row precede its writes in program order, so one barrier per
// column suffices.  No rows are swapped: pivots are recorded by their LDS
// row, and the identity columns index LDS rows directly.

// A value held alike in every lane, moved to scalar registers (so that code
// depending on it alone runs on the scalar unit).
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// v as held by lane l (l uniform), in scalar registers
__device__ __forceinline__ uint32_t lane_value(uint32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ uint64_t lane_value(uint64_t v, int l) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t uniform(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

template <class LT, bool CMP, int IF_T>
__device__ static void inner_factor_body(typename LT::T* __restrict__ X, int64_t N,
                                         uint8_t* __restrict__ used, int32_t* __restrict__ slot_of,
                                         int32_t* __restrict__ epoch_of, int epoch,
                                         int64_t colbase, typename LT::F f, ElimState* st,
                                         int64_t* __restrict__ piv_rows, int64_t* __restrict__ piv_cols,
                                         BlockState<typename LT::T>* bs) {
    using T = typename LT::T;
    constexpr int IF_RPW = CI / (IF_T / 64);  // candidate rows per wave
    static_assert(IF_RPW * (IF_T / 64) == CI && IF_T >= 256, "waves split the candidate rows evenly; SCAN rows per probe");
    // CMP: the compact form (32 slots a row), else [panel | identity] (64
    // columns a row); one padding column
    static_assert(!CMP || IF_RPW % 2 == 0, "the compact form takes rows in pairs");
    constexpr int AW = (CMP ? BI : BI + CI) + 1;
    __shared__ T aug[CI][AW];
    __shared__ int64_t cand[CI + 1];
    __shared__ int prow_of[CI];   // LDS row of the t-th pivot
    __shared__ int pcol[CI];      // its column within the inner panel
    __shared__ int s_total;
    __shared__ int s_wcnt[IF_T / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const unsigned long long t0 = wall_clock64();
    const int c0 = (int)uniform((uint32_t)bs->c0_next);
    const int bcur = (int)uniform((uint32_t)bs->bcur);
    if (c0 >= bcur) {  // every inner panel of this outer panel is complete
        if (tid == 0) bs->k = 0;
        return;
    }
    const int w = min(BI, bcur - c0);

    if (tid == 0) s_total = 0;
    __syncthreads();
    // 1. the first CI (+1 to know whether more exist) candidate rows, in row
    //    order, probing SCAN rows at a time with the first SCAN / 64 waves
    //    (larger probes measured slower: most steps need the first few dozen)
    constexpr int SCAN = 256;
    for (int64_t base = st->first_unused; base < N; base += SCAN) {
        // unused row with a non-zero entry in the inner panel; the flag and the
        // row are fetched together (rows past N read row N-1 and are dropped)
        const int64_t r = base + tid, rc = r < N ? r : N - 1;
        bool fl = false;
        if (tid < SCAN) {
            constexpr int PER = 16 / (int)sizeof(T);  // elements per 16-byte vector
            const Vec16<T>* row = reinterpret_cast<const Vec16<T>*>(X + rc * XLD + c0);  // 128-byte aligned
            const uint8_t u = gld<uint8_t>(used + rc);
            Vec16<T> v[BI / PER];
#pragma unroll
            for (int q = 0; q < BI / PER; ++q) {
                const v4i raw = gld<v4i>(row + q);
                __builtin_memcpy(&v[q], &raw, 16);
            }
            T acc = 0;
#pragma unroll
            for (int q = 0; q < BI / PER; ++q)
#pragma unroll
                for (int e = 0; e < PER; ++e) acc |= (PER * q + e < w ? v[q].v[e] : (T)0);
            fl = r < N && !u && acc != (T)0;
        }
        const unsigned long long m = __ballot(fl);
        if (lane == 0 && tid < SCAN) s_wcnt[wid] = __popcll(m);
        __syncthreads();
        if (tid < SCAN) {
            int before = s_total;
            for (int q = 0; q < wid; ++q) before += s_wcnt[q];
            const int pos = before + __popcll(m & ((1ull << lane) - 1ull));
            if (fl && pos <= CI) cand[pos] = r;
        }
        __syncthreads();
        if (tid == 0) {
            int sum = 0;
            for (int q = 0; q < SCAN / 64; ++q) sum += s_wcnt[q];
            s_total += sum;
        }
        __syncthreads();
        if (s_total > CI) break;
    }
    const unsigned long long t1 = wall_clock64();
    // (read from LDS, so uniform in fact but not to the compiler: in scalar
    // registers the elimination loop below is scalar control flow)
    const int total = (int)uniform((uint32_t)s_total);
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

    // 2. the candidate rows in LDS: [X[S, c0:c0+w] | I] in the 32-bit form,
    //    X[S, c0:c0+w] alone in the compact 64-bit form (below); rows >= nC
    //    and columns >= w are zero and never pivot
    {
        const int j = lane;
#pragma unroll
        for (int q = 0; q < IF_RPW; ++q) {
            const int ii = wid + (IF_T / 64) * q;
            T v = 0;
            if (ii < nC) {
                if (j < w) v = X[cand[ii] * XLD + c0 + j];
                else if (!CMP && j >= BI) v = (j - BI == ii) ? 1 : 0;
            }
            if (j < AW - 1) aug[ii][j] = v;
        }
    }
    __syncthreads();

    const unsigned long long t2 = wall_clock64();
    // 3. fraction-free Gauss-Jordan: row_i <- (h * row_i - a_i * row_pv) R^{-1}
    //    (rows stay non-zero multiples of the exact Gauss-Jordan rows).  The
    //    pivot heads h_t are kept: the pivot of the t-th pivot row ends as
    //    d_t = h_t prod_{u > t} h_u R^{-1}, which step 4 inverts.
    unsigned long long pmask = 0ull;  // LDS rows already used as pivots (same in every wave)
    const unsigned long long live = (1ull << nC) - 1ull;  // the candidate rows (nC <= CI < 64)
    int t = 0;
    constexpr int NW = IF_T / 64;  // wave w owns rows w, w + NW, w + 2 NW, ...
    __shared__ T s_h[CI];
    if constexpr (!CMP) {
        for (int c = 0; c < w && t < nC; ++c) {
            // the wave's own rows are read beside the pivot search (all reads
            // of them come before any write)
            T av[IF_RPW], xv[IF_RPW];
#pragma unroll
            for (int q = 0; q < IF_RPW; ++q) {
                const int r = wid + NW * q;
                av[q] = aug[r][c];
                xv[q] = aug[r][lane];
            }
            // the pivot is the first row not yet a pivot with a non-zero entry:
            // the first row not yet a pivot (ps) whenever its entry is non-zero,
            // which is read beside the rows above; the search over all rows
            // (a second round trip to LDS) only when it is zero
            const int ps = __ffsll((long long)(~pmask & live)) - 1;
            T y = aug[ps][lane];
            T h = lane_value(y, c);  // its entry in column c, from lane c (no second load)
            int pv = ps;
            if (h == (T)0) {
                const bool nz = lane < nC && !((pmask >> lane) & 1ull) && aug[lane < CI ? lane : 0][c] != (T)0;
                const unsigned long long m = __ballot(nz);
                if (m == 0ull) continue;  // no pivot in this column (every wave sees the same)
                pv = __ffsll((long long)m) - 1;
                h = aug[pv][c];
                y = aug[pv][lane];
            }
            // every row's update first (independent chains, interleaved), then
            // the stores of the rows that change
            T nv[IF_RPW];
            f.template ff_update_n<IF_RPW>(h, xv, av, y, nv);
#pragma unroll
            for (int q = 0; q < IF_RPW; ++q) {
                const int r = wid + NW * q;
                if (r < nC && r != pv) aug[r][lane] = nv[q];
            }
            if (tid == 0) {
                prow_of[t] = pv;
                pcol[t] = c;
                s_h[t] = h;
            }
            pmask |= 1ull << pv;
            ++t;
            __syncthreads();
        }
    } else {
        // Compact form: 32 slots per row, two rows per wave instruction (lanes
        // 32 hf + j, slot j).  Slot j holds panel column j until column j is
        // eliminated; from then on it holds the identity-part column of the
        // row that pivoted there, scaled by a per-column constant.  Before step
        // c every row that is not yet a pivot carries D_c = prod_{u < t} h_u
        // R^{-1} on its own identity column and nothing on the others', so
        // the identity part has at most t live columns besides that one
        // diagonal, and t <= c slots are free to hold them.  Step c with pivot
        // row pv writes the pivot's identity column, (0 h - a_i D_c) R^{-1} in
        // row i and D_c in row pv, as -a_i and R (mod p) with the scale
        // s_c = D_c R^{-1}, into slot c, where the panel column has just become
        // zero (h in row pv).  Half the lanes of the 64-wide form, whose
        // 64-bit products run at a quarter of the VALU rate.  The write into
        // row pv waits until after the barrier (other waves read h there in
        // this step); its owner lane makes it before it reads the row again.
        const T one_m = f.to_mont((T)1);
        const int hf = lane >> 5, j = lane & 31;
        constexpr int RP = IF_RPW / 2;  // row pairs per wave: rows wid + NW (2 q + hf)
        int pend = -1;                  // LDS row whose slot pend_c takes R (mod p)
        int pend_c = 0;
        for (int c = 0; c < w && t < nC; ++c) {
            if (pend >= 0 && wid == pend % NW && lane == 32 * ((pend / NW) & 1) + pend_c) aug[pend][pend_c] = one_m;
            pend = -1;
            T av[RP], xv[RP];
#pragma unroll
            for (int q = 0; q < RP; ++q) {
                const int r = wid + NW * (2 * q + hf);
                av[q] = aug[r][c];
                xv[q] = aug[r][j];
            }
            // (the first row not yet a pivot, unless its entry is zero: as above)
            const int ps = __ffsll((long long)(~pmask & live)) - 1;
            T y = aug[ps][j];
            T h = lane_value(y, c);
            int pv = ps;
            if (h == (T)0) {
                const bool nz = lane < nC && !((pmask >> lane) & 1ull) && aug[lane < CI ? lane : 0][c] != (T)0;
                const unsigned long long m = __ballot(nz);
                if (m == 0ull) continue;
                pv = __ffsll((long long)m) - 1;
                h = aug[pv][c];
                y = aug[pv][j];
            }
            T nv[RP];
            f.template ff_update_n<RP>(h, xv, av, y, nv);
#pragma unroll
            for (int q = 0; q < RP; ++q)
                if (j == c) nv[q] = av[q] != (T)0 ? f.p - av[q] : (T)0;
#pragma unroll
            for (int q = 0; q < RP; ++q) {
                const int r = wid + NW * (2 * q + hf);
                if (r < nC && r != pv) aug[r][j] = nv[q];
            }
            if (tid == 0) {
                prow_of[t] = pv;
                pcol[t] = c;
                s_h[t] = h;
            }
            pmask |= 1ull << pv;
            pend = pv;
            pend_c = c;
            ++t;
            __syncthreads();
        }
        if (pend >= 0 && wid == pend % NW && lane == 32 * ((pend / NW) & 1) + pend_c) aug[pend][pend_c] = one_m;
    }
    const int k = t;
    __syncthreads();
    const unsigned long long t3 = wall_clock64();

    // 4. the pivots' inverses from one inversion: with P_t = prod_{u < t} h_u
    //    R^{-(t-1)} (P_0 = R), an exclusive product scan of the heads over the
    //    lanes of wave 0, d_t^{-1} = P_t R^{-1} / P_k.  P_k is the same in
    //    every lane, so its inverse runs once, on the scalar unit.
    //    s_inv[t] = d_t^{-1} R, s_pre[t] = P_t.
    __shared__ T s_inv[CI], s_pre[CI];
    if (wid == 0 && k > 0) {
        const T one_m = f.to_mont((T)1);
        T v = lane < k ? s_h[lane] : one_m;
#pragma unroll
        for (int off = 1; off < CI; off <<= 1) {
            const T o = __shfl_up(v, off, 64);
            if (lane >= off) v = f.mont(o, v);
        }
        T ex = __shfl_up(v, 1, 64);
        if (lane == 0) ex = one_m;
        const T tot = uniform(__shfl(v, k - 1, 64));
        T inv;
        if constexpr (sizeof(T) == 4) inv = euclid_inverse_u32(tot, f.p);
        else inv = f.m61 ? m61_inverse(tot) : binary_inverse<T>(tot, f.p);
        const T x = f.to_mont(inv);
        if (lane < k) {
            s_pre[lane] = ex;
            s_inv[lane] = f.mont(ex, x);
        }
    }
    __syncthreads();

    const unsigned long long t4 = wall_clock64();
    // 5. outputs: Apiv^{-1} (Montgomery form), pivot rows/columns, bookkeeping.
    //    Entry (a, b) is the identity-part entry of pivot row a in column b
    //    over d_a (times R for the Montgomery form); the compact form also
    //    applies the column's scale s_b = P_b R^{-2}.
    const int64_t rank0 = st->rank;
    const int q0 = bs->q0;
    for (int x = tid; x < k * k; x += IF_T) {
        const int a = x / k, b = x - a * k;
        T out;
        if constexpr (CMP) out = f.mont(f.mont(aug[prow_of[a]][pcol[b]], s_pre[b]), s_inv[a]);
        else out = f.to_mont(f.mont(s_inv[a], aug[prow_of[a]][BI + prow_of[b]]));
        bs->apinv[a * CI + b] = out;
    }
    if (tid < k) {
        const int64_t r = cand[prow_of[tid]];
        bs->R[tid] = r;
        bs->Cp[tid] = pcol[tid];
        used[r] = 1;
        epoch_of[r] = epoch;
        slot_of[r] = q0 + tid;  // its slot in the outer panel (the column of Z holding its 1)
        piv_rows[rank0 + tid] = r;
        piv_cols[rank0 + tid] = colbase + c0 + pcol[tid];
        X[r * XLD + BO + q0 + tid] = 1;  // Z[r, slot] = 1
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


// IF_T threads (LT::IF_T and its register cap, or fewer waves, each with more
// candidate rows, capped only by what one workgroup needs)
template <class LT, bool CMP, int IF_T>
__global__ __launch_bounds__(IF_T, IF_T == LT::IF_T ? LT::IF_WAVES : IF_T / 256) void k_inner_factor(
    const BMat<LT>* __restrict__ D, int epoch) {
    const BMat<LT>& d = D[blockIdx.x];
    inner_factor_body<LT, CMP, IF_T>(d.X, d.N, d.used, d.slot_of, d.epoch_of, epoch, d.J, d.f, d.st, d.piv_rows,
                                     d.piv_cols, d.bs);
}

// The inner factorisation of storage width LT: the compact form always for
// 64-bit storage, by tuning knob "inner_compact" for 32-bit storage; the
// workgroup size from tuning knob "inner_threads" (0: LT::IF_T).
template <class LT>
static void launch_inner_factor(unsigned B, hipStream_t s, const BMat<LT>* desc, int epoch) {
    if constexpr (std::is_same<LT, LimbsU64>::value) {
        switch (tuning().inner_threads) {
            case 256: k_inner_factor<LT, true, 256><<<B, 256, 0, s>>>(desc, epoch); break;
            case 1024: k_inner_factor<LT, true, 1024><<<B, 1024, 0, s>>>(desc, epoch); break;
            default: k_inner_factor<LT, true, LT::IF_T><<<B, LT::IF_T, 0, s>>>(desc, epoch);
        }
    } else {
        const bool cmp = tuning().inner_compact != 0;
        switch (tuning().inner_threads) {
            case 256:
                if (cmp) k_inner_factor<LT, true, 256><<<B, 256, 0, s>>>(desc, epoch);
                else k_inner_factor<LT, false, 256><<<B, 256, 0, s>>>(desc, epoch);
                break;
            case 512:
                if (cmp) k_inner_factor<LT, true, 512><<<B, 512, 0, s>>>(desc, epoch);
                else k_inner_factor<LT, false, 512><<<B, 512, 0, s>>>(desc, epoch);
                break;
            default:
                if (cmp) k_inner_factor<LT, true, LT::IF_T><<<B, LT::IF_T, 0, s>>>(desc, epoch);
                else k_inner_factor<LT, false, LT::IF_T><<<B, LT::IF_T, 0, s>>>(desc, epoch);
        }
    }
}

// U = Apiv^{-1} X[R, :] (k x XLD, Montgomery inverse times plain rows gives plain
// values) and the limb planes of both operands of the inner GEMM:
//   planesU[l][n][t] = digit l of U[t, n]                (t < k, zero up to KPI)
//   planesA[l][r][t] = digit l of X[r, c0 + Cp[t]]       (every row r; 64-bit only)
// The inner GEMM overwrites the pivot columns X[:, c0 + Cp] in place, so when
// several of its workgroups share a row stripe (always for 64-bit storage) its
// left operand is captured here, before any of them runs; a 32-bit workgroup
// that owns its stripe captures its own.  The pivot rows X[R, :] get U here.
// Workgroups [0, MU_BLOCKS) build U (2 columns each; each entry's dot product
// of length k split over MU_SPLIT adjacent lanes and summed by shuffles, so
// that its chain of dependent products is k / MU_SPLIT long), the rest
// planesA (MU_ROWS rows each, grid-strided).
constexpr int MU_COLS = 2;
constexpr int MU_SPLIT = 4;
constexpr int MU_BLOCKS = XLD / MU_COLS;
static_assert(MU_COLS * CI * MU_SPLIT == 256 && 2 * CI == KPI, "one thread per (column, pivot slot, part); slots pad to KPI");
constexpr int MU_ROWS = 16;
// Column pairs per U workgroup: every U workgroup reads the whole k x k
// inverse, so a batch of many matrices takes several pairs per workgroup
// (tuning knob "u_passes"; 0: 1 below 8 matrices, 4 from 8 on).
static inline int u_passes(unsigned B) {
    const int64_t k = tuning().u_passes;
    return k > 0 ? (int)k : (B >= 8 ? 4 : 1);
}

template <class LT>
__device__ static void make_u_body(typename LT::T* __restrict__ X, int64_t N, typename LT::F f,
                                   const BlockState<typename LT::T>* __restrict__ bs, typename LT::T* __restrict__ U,
                                   int8_t* __restrict__ planesU, int64_t strideU,
                                   int8_t* __restrict__ planesA, int64_t strideA, int passes) {
    using T = typename LT::T;
    constexpr int NL = LT::NL;
    __shared__ T s_inv[CI * CI];
    __shared__ int64_t s_R[CI];
    __shared__ int s_cp[CI];
    const int k = bs->k;
    if (k == 0) return;  // the inner GEMM skips steps without pivots
    const int t = threadIdx.x;  // 0..63
    const int ublocks = MU_BLOCKS / passes;  // U workgroups, `passes` column pairs each
    if ((int)blockIdx.x >= ublocks) {
        const int c0 = bs->c0_cur;
        if (threadIdx.y == 0 && t < k) s_cp[t] = bs->Cp[t];
        __syncthreads();
        const int64_t col = t < k ? c0 + s_cp[t] : 0;
        for (int64_t r0 = (int64_t)(blockIdx.x - ublocks) * MU_ROWS + threadIdx.y; r0 < N;
             r0 += (int64_t)(gridDim.x - ublocks) * MU_ROWS) {
            T v[MU_ROWS / 4];
#pragma unroll
            for (int i = 0; i < MU_ROWS / 4; ++i) {
                const int64_t r = r0 + 4 * i;
                v[i] = (t < k && r < N) ? X[r * XLD + col] : (T)0;
            }
#pragma unroll
            for (int i = 0; i < MU_ROWS / 4; ++i) {
                const int64_t r = r0 + 4 * i;
                if (r >= N) break;
                int8_t a[NL] = {};
                if (t < k) balanced_digits(centred(v[i], f.p), a);
#pragma unroll
                for (int l = 0; l < NL; ++l) planesA[l * strideA + r * KPI + t] = a[l];
            }
        }
        return;
    }
    const int fl = threadIdx.x + threadIdx.y * 64;
    for (int x = fl; x < k * CI; x += 256) s_inv[x] = bs->apinv[x];
    if (fl < k) s_R[fl] = bs->R[fl];
    __syncthreads();
    // the block's columns of the k pivot rows, fetched at once (a load per
    // term of the dot product would expose the memory latency k times)
    constexpr int MU_PASSES_MAX = 8;
    __shared__ T s_x[CI][MU_COLS * MU_PASSES_MAX];
    const int ncols = MU_COLS * passes;
    for (int x = fl; x < k * ncols; x += 256) {
        const int sl = x / ncols, yy = x - sl * ncols;
        s_x[sl][yy] = X[s_R[sl] * XLD + (int64_t)blockIdx.x * ncols + yy];
    }
    __syncthreads();
    // lane layout: part = fl % MU_SPLIT (adjacent lanes), pivot slot, column
    const int part = fl % MU_SPLIT, ts = (fl / MU_SPLIT) % CI, y0 = fl / (MU_SPLIT * CI);
    auto one_pass = [&](int y) {
        const int64_t n = (int64_t)blockIdx.x * ncols + y;
        T u = 0;
        if (ts < k)
            for (int s = part; s < k; s += MU_SPLIT) u = f.add(u, f.mont(s_inv[ts * CI + s], s_x[s][y]));
#pragma unroll
        for (int off = 1; off < MU_SPLIT; off <<= 1) u = f.add(u, (T)__shfl_xor(u, off, 64));
        if (ts < k && part == 0) {
            U[ts * XLD + n] = u;
            // the pivot rows' new values go straight into the panel (the inner
            // GEMM leaves those rows alone); every read of X by this workgroup is
            // behind the barrier above, and the planes workgroups' reads of pivot
            // rows feed only results that are never stored
            X[s_R[ts] * XLD + n] = u;
        }
        int8_t d[NL] = {};
        if (ts < k) balanced_digits(centred(u, f.p), d);
        // part 0 writes slot ts, part 1 slot ts + CI (the zero padding up to KPI)
        if (part < 2) {
            const int slot = ts + part * CI;
#pragma unroll
            for (int l = 0; l < NL; ++l) planesU[l * strideU + n * KPI + slot] = part == 0 ? d[l] : (int8_t)0;
        }
    };
    if (passes == 1) one_pass(y0);
    else
        for (int pass = 0; pass < passes; ++pass) one_pass(y0 + MU_COLS * pass);
}


template <class LT>
__global__ __launch_bounds__(256) void k_make_u(const BMat<LT>* __restrict__ D, int passes) {
    const BMat<LT>& d = D[blockIdx.y];
    make_u_body<LT>(d.X, d.N, d.f, d.bs, d.U, d.planesUi, (int64_t)XLD * KPI, d.planesAi, d.strideAi, passes);
}

// ---------------------------------------------------------------------------
// outer update operands
// ---------------------------------------------------------------------------

// The trailing update in the form P = I + W S: S selects the panel's pivot
// rows Q and W = Z - E, where E has a 1 at (Q[i], i), so that one product
//   A[:, T] <- A[:, T] + W A[Q, T]
// also gives the pivot rows their new values Z[Q] A[Q, T].  Two consecutive
// panels compose into P' P = I + [W + W' W[Q'] | W'] [S; S'], so the update of
// the columns beyond the next panel is deferred by one panel and then done as
// one GEMM of depth |Q| + |Q'| <= 512 (half the folds and C traffic per MAC).
// K layout of both operands: [deferred part, KqA = kA rounded up to 64 |
// current panel, q0 rounded up to 64], zero padded.

// Limb planes of -W, the left operand, for one matrix (workgroup bx of nbx
// striding over rows; the 1 of E in pivot row m sits in column slot_of[m]):
//   src 0: the current panel, W[m, t] = Z[m, t] - [m = Q[t]] (Z = X[:, BO + t], t < q0),
//          written at K offset KqA; kept in d.W when the panel's update is deferred;
//   src 1: the deferred panel's composite transformation d.W[:, t] (t < kA), at K offset 0.
template <class LT>
__device__ static void make_lhs_body(const BMat<LT>& d, int src, int e1, int64_t bx, int64_t nbx) {
    using T = typename LT::T;
    constexpr int NL = LT::NL;
    if (d.nT == 0 && !(src == 0 && d.save_w)) return;
    const int width = src == 0 ? d.q0 : d.kA;
    const int kw = src == 0 ? d.Kq - d.KqA : d.KqA;  // a multiple of 64, at most BO
    const int koff = src == 0 ? d.KqA : 0;
    if (width == 0) return;
    const bool save = src == 0 && d.save_w;
    // four consecutive columns t per thread, BO / 4 threads per row
    const int t = (threadIdx.x & (BO / 4 - 1)) * 4;
    if (t >= kw) return;
    constexpr int RPB = 256 / (BO / 4);  // rows per block and pass
    for (int64_t m = d.r0 + bx * RPB + threadIdx.x / (BO / 4); m < d.N; m += nbx * RPB) {
        int8_t dg[NL][4] = {};
        if (t < width) {
            T w[4];
            if (src == 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e) w[e] = d.X[m * XLD + BO + t + e];
                const int32_t ep = d.epoch_of[m];
                if (ep >= d.e0 && ep < e1) {
                    const int sl = d.slot_of[m] - t;
                    if (sl >= 0 && sl < 4) w[sl] = d.f.sub(w[sl], (T)1);
                }
                if (save)
#pragma unroll
                    for (int e = 0; e < 4; ++e) d.W[m * BO + t + e] = w[e];
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) w[e] = d.W[m * BO + t + e];
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int8_t g[NL];
                balanced_digits(t + e < width ? -centred(w[e], d.f.p) : 0, g);
#pragma unroll
                for (int l = 0; l < NL; ++l) dg[l][e] = g[l];
            }
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            int32_t pk;
            __builtin_memcpy(&pk, dg[l], 4);
            *reinterpret_cast<int32_t*>(d.planesL + l * d.strideL + ptile(m, koff + t, d.Kq)) = pk;
        }
    }
}

// Limb planes of the right operand, transposed, for one matrix (workgroup bx
// of nbx striding over column tiles, by = 64-deep K tile), 64 x 64 tiles
// through LDS so that both the row reads and the plane writes are coalesced.
//   job 0: planesB[l][n][t] = digit l of A[row(t), T0 + n], row(t) = the
//          deferred panel's pivot Q_A[t] (t < kA) or the current panel's
//          Q[t - KqA] (t - KqA < q0); zero elsewhere.  Row stride Kq.
//   job 1: planesW[l][n][t] = digit l of W[Q[t], n] (n < kA, t < q0): the
//          factor W[Q'] of the composite transformation.  Row stride Kq - KqA.
template <class LT>
__device__ static void make_rhs_body(const BMat<LT>& d, int job, int64_t bx, int64_t nbx, int by) {
    using S = typename LT::S;
    constexpr int NL = LT::NL;
    __shared__ S tile[64][65];
    const int64_t ncols = job == 0 ? d.nT : (d.q0 > 0 ? d.kA : 0);
    const int depth = job == 0 ? d.Kq : d.Kq - d.KqA;
    if (ncols == 0) return;
    const int tb = by * 64;
    if (tb >= depth) return;
    const int64_t* Q = d.piv_rows + d.rank_start;
    const int64_t* QA = Q - d.kA;
    int8_t* planes = job == 0 ? d.planesB : d.planesW;
    const int64_t pstride = job == 0 ? d.strideB : STRIDE_W;
    for (int64_t nb = bx * 64; nb < ncols; nb += nbx * 64) {
        // all sixteen loads of a thread in flight together (one after the
        // other they expose the memory latency sixteen times)
        typename LT::T raw[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int x = threadIdx.x + 256 * u;
            const int tt = x >> 6, nn = x & 63;
            const int t = tb + tt;
            const int64_t n = nb + nn;
            const typename LT::T* src = nullptr;
            if (n < ncols) {
                if (job == 1) {
                    if (t < d.q0) src = d.W + Q[t] * BO + n;
                } else if (t < d.KqA) {
                    if (t < d.kA) src = d.A + QA[t] * d.ld + d.T0 + n;
                } else if (t - d.KqA < d.q0) {
                    src = d.A + Q[t - d.KqA] * d.ld + d.T0 + n;
                }
            }
            raw[u] = src ? *src : (typename LT::T)0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int x = threadIdx.x + 256 * u;
            tile[x >> 6][x & 63] = centred(raw[u], d.f.p);
        }
        __syncthreads();
        const int nn = threadIdx.x >> 2, tq = (threadIdx.x & 3) * 16;
        const int64_t n = nb + nn;
        if (n < ncols) {
            // 16 digits of each limb packed into four words with shifts (a
            // byte array written element-wise ends up in scratch memory)
            uint32_t wd[NL][4];
#pragma unroll
            for (int l = 0; l < NL; ++l)
#pragma unroll
                for (int q = 0; q < 4; ++q) wd[l][q] = 0u;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                int8_t dg[NL];
                balanced_digits(tile[tq + i][nn], dg);
#pragma unroll
                for (int l = 0; l < NL; ++l) wd[l][i >> 2] |= (uint32_t)(uint8_t)dg[l] << (8 * (i & 3));
            }
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                const v4i w = {(int)wd[l][0], (int)wd[l][1], (int)wd[l][2], (int)wd[l][3]};
                *reinterpret_cast<v4i*>(planes + l * pstride + ptile(n, tb + tq, depth)) = w;
            }
        }
        __syncthreads();
    }
}

template <class LT>
__global__ __launch_bounds__(256) void k_make_rhs(const BMat<LT>* __restrict__ D, int job) {
    make_rhs_body<LT>(D[blockIdx.z], job, blockIdx.x, gridDim.x, blockIdx.y);
}

// Both operands' planes in one launch (blockIdx.y = matrix): workgroups
// [0, nl) build the left operand (source lsrc, make_lhs_body; none if
// nl = 0), the rest the right operand of job rjob in rx x ry tiles
// (make_rhs_body; none if rjob < 0).  The two read and write disjoint data.
template <class LT>
__global__ __launch_bounds__(256) void k_make_ops(const BMat<LT>* __restrict__ D, int lsrc, int e1, int nl, int rjob,
                                                  int rx) {
    const BMat<LT>& d = D[blockIdx.y];
    if ((int)blockIdx.x < nl) {
        make_lhs_body<LT>(d, lsrc, e1, blockIdx.x, nl);
    } else if (rjob >= 0) {
        const int r = (int)blockIdx.x - nl;
        make_rhs_body<LT>(d, rjob, r % rx, rx, r / rx);
    }
}

// ---------------------------------------------------------------------------
// modular GEMMs on the int8 matrix cores (see modp_gemm.hpp for the limb
// arithmetic): 256 threads, 64 x 64 tile (4 waves of 32 x 32), 64-deep K
// stages double-buffered in LDS (two workgroups per CU: one's prologue and
// epilogue overlap the other's MFMAs).  The flattened tile list of the batch
// is cut into 8 contiguous ranges, one per XCD (consecutive workgroup ids go
// to different XCDs), and each matrix's tiles are walked in groups of tile
// rows (tuning knob "gemm_group", default 8) so that a B tile is reused that
// many times while the A tiles stay in L2.
// ---------------------------------------------------------------------------

constexpr int G_M = 64, G_N = 64, G_K = 64, G_T = 256;
constexpr int G_ROW = G_K;  // LDS bytes per staged row (no padding: the 16-byte columns are swizzled)

// LDS offset of the 16-byte column `qu` of staged row `row`.  The XOR with
// bits 2-3 of the row makes both access patterns conflict-free: the MFMA
// fragment reads (16 consecutive rows, one column) and the stage stores
// (4 consecutive rows, all 4 columns, per 16 lanes).
__device__ __forceinline__ int lds_at(int row, int qu) { return row * G_ROW + ((qu ^ ((row >> 2) & 3)) << 4); }
constexpr int NUM_XCD = 8;

// blockIdx.x -> (matrix, tile row, tile column) through the prefix sums of
// per-matrix tile counts `pre[0..B]` (row-tile counts in `mts`).
struct TileSched {
    const int64_t* pre;  // [B + 1]
    const int64_t* mts;  // [B]
    int B;
    int group;           // tile rows walked together (tuning knob "gemm_group")
    int64_t t0, t1;      // the launch covers tiles [t0, t1) of the flattened list (t1 < 0: to its end)
};

__device__ __forceinline__ bool batch_tile(const TileSched& ts, int* m_out, int64_t* tm, int64_t* tn) {
    const int64_t* __restrict__ pre = ts.pre;
    const int64_t* __restrict__ mts = ts.mts;
    const int B = ts.B, group = ts.group;
    const int64_t end = ts.t1 < 0 ? pre[B] : ts.t1;
    const int64_t per_xcd = (end - ts.t0 + NUM_XCD - 1) / NUM_XCD;
    const int64_t bid = blockIdx.x;
    // (a grid larger than the schedule -- the fast path's, sized before its
    // gate drops matrices -- must not reach into the next XCD's range)
    if (bid / NUM_XCD >= per_xcd) return false;
    const int64_t t = ts.t0 + (bid % NUM_XCD) * per_xcd + bid / NUM_XCD;
    if (t >= end) return false;
    int lo = 0, hi = B;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= t) lo = mid;
        else hi = mid;
    }
    // tile counts stay far below 2^32: 32-bit divisions (a 64-bit one costs
    // several times as many instructions, once per workgroup)
    const uint32_t loc = (uint32_t)(t - pre[lo]);
    const uint32_t mt = (uint32_t)mts[lo];
    const uint32_t nt = (uint32_t)(pre[lo + 1] - pre[lo]) / mt;
    const uint32_t gsz = (uint32_t)group * nt;
    const uint32_t g = loc / gsz, tt = loc - g * gsz;
    const uint32_t gm = min((uint32_t)group, mt - g * (uint32_t)group);
    *m_out = lo;
    *tm = (int64_t)(g * (uint32_t)group + tt % gm);
    *tn = (int64_t)(tt / gm);
    return true;
}

// Trailing-update GEMMs, C <- C - L B with L = limbs of -W (see make_lhs_body):
//   JOB 0: A[:, T] <- A[:, T] + W A[Qc, T]   (the whole depth Kq, every row);
//   JOB 1: W <- W + W' W[Q']                 (the composite transformation of a
//          deferred panel: L = the current panel's part, B = planesW).
// This form stages through registers; k_gemm_outer_lds (below) does the same
// work with direct-to-LDS loads and runs wherever C rows are 16-byte aligned.
template <int JOB>
__global__ __launch_bounds__(G_T, 2) void k_gemm_outer(const BMat<LimbsU32>* __restrict__ D, TileSched ts) {
    __shared__ __attribute__((aligned(16))) int8_t sA[2][4][G_M * G_ROW];
    __shared__ __attribute__((aligned(16))) int8_t sB[2][4][G_N * G_ROW];
    int mi_;
    int64_t tm, tn;
    if (!batch_tile(ts, &mi_, &tm, &tn)) return;
    const BMat<LimbsU32>& d = D[mi_];
    uint32_t* C = JOB == 0 ? d.A + d.T0 : d.W;
    const int64_t ldc = JOB == 0 ? d.ld : BO;
    const int64_t Mrows = d.N;
    const int64_t Ncols = JOB == 0 ? d.nT : d.kA;
    const int64_t ldL = d.Kq;
    const int8_t* L = d.planesL + (JOB == 0 ? 0 : 16 * (int64_t)d.KqA);
    const int64_t planeL = d.strideL;
    const int8_t* B = JOB == 0 ? d.planesB : d.planesW;
    const int64_t planeB = JOB == 0 ? d.strideB : STRIDE_W;
    const int Kpad = JOB == 0 ? d.Kq : d.Kq - d.KqA;
    const double pd = d.pd, pinv = d.pinv;
    const int64_t m_base = d.r0 + tm * G_M, n_base = tn * G_N;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int r16 = lane & 15, g = lane >> 4;
    const v4i zero = {0, 0, 0, 0};

    // Loads are unconditional: rows and columns beyond the matrix read a
    // clamped (valid) address, and the accumulators they feed are never
    // stored.  (A select or branch around a load makes the compiler wait for
    // it where the value is merged, not where it is used.)
    // C values of this lane, fetched while the last K stage runs on the matrix
    // cores (the stage registers are free by then)
    uint32_t cpre[2][2][4];
    auto cload = [&]() {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int64_t row = min(m_base + wm * 32 + mi * 16 + 4 * g + i, Mrows - 1);
                    const int64_t col = min(n_base + wn * 32 + ni * 16 + r16, Ncols - 1);
                    cpre[mi][ni][i] = gld<uint32_t>(C + row * ldc + col);
                }
    };
    // stage copy: A and B = 4 limbs x 64 rows x 4 sixteen-byte quarters (4
    // each per thread); the row/limb part of each address is fixed per tile
    v4i ra[4], rb[4];
    const int srow = tid >> 2, squ = tid & 3;  // the same for the 4 limbs (q = limb)
    const int8_t* pa = L + ptile(min(m_base + srow, Mrows - 1), squ * 16, ldL);
    const int8_t* pb = B + ptile(min(n_base + srow, Ncols - 1), squ * 16, Kpad);
    const int soff = lds_at(srow, squ);
    auto gload = [&](int kc) {  // (depth kc: 16 kc bytes into the tiled planes)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ra[q] = gld<v4i>(pa + q * planeL + 16 * kc);
            rb[q] = gld<v4i>(pb + q * planeB + 16 * kc);
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            *reinterpret_cast<v4i*>(&sA[buf][q][soff]) = ra[q];
            *reinterpret_cast<v4i*>(&sB[buf][q][soff]) = rb[q];
        }
    };

    v4i acc[7][2][2];
#pragma unroll
    for (int s = 0; s < 7; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[s][i][j] = zero;

    const int nk = Kpad / G_K;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kk = 0; kk < nk; ++kk) {
        const int buf = kk & 1;
        if (kk + 1 < nk) gload((kk + 1) * G_K);
        else cload();
        v4i bf[4][2];
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
                bf[l][ni] = *reinterpret_cast<const v4i*>(&sB[buf][l][lds_at(wn * 32 + ni * 16 + r16, g)]);
#pragma unroll
        for (int l1 = 0; l1 < 4; ++l1) {
            v4i af[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
                af[mi] = *reinterpret_cast<const v4i*>(&sA[buf][l1][lds_at(wm * 32 + mi * 16 + r16, g)]);
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
                    int32_t sv[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) sv[q] = acc[q][mi][ni][i];
                    C[row * ldc + col] = fold_sub(cpre[mi][ni][i], sv, pd, pinv);
                }
            }
}

// ---- direct-to-LDS staging (gfx950 global_load_lds_dwordx4) ----
// s_waitcnt vmcnt(N) with the other counters left alone.
template <int N>
__device__ __forceinline__ void vm_wait() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}
// 16 bytes per lane from global memory straight into LDS: lane i writes
// bytes [dst + 16 i, dst + 16 i + 16) (dst uniform across the wave).  Written
// as asm so that the compiler does not track it (its own tracking would wait
// for every such load before each LDS read); callers wait with
// lds_dma_barrier.  The asm has no register outputs.
__device__ __forceinline__ void dma16(const void* src, void* dst) {
    const uint32_t lds =
        __builtin_amdgcn_readfirstlane((uint32_t)(size_t)(__attribute__((address_space(3))) int8_t*)dst);
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(lds) : "memory", "m0");
}
// The same with the address split the instruction's way: a wave-uniform
// 64-bit base in scalar registers plus a 32-bit per-lane offset, and the LDS
// destination given as a wave-uniform LDS byte address (no vector address
// arithmetic per load: a stage's bases advance on the scalar unit).
__device__ __forceinline__ void dma16_so(uint32_t voff, const void* sbase, uint32_t lds) {
    asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(voff), "s"(sbase), "s"(lds)
                 : "memory", "m0");
}
// every wave's direct-to-LDS loads have landed and every wave is past its
// reads of the buffers they refill
__device__ __forceinline__ void lds_dma_barrier() {
    vm_wait<0>();
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}
typedef int8_t StageBuf[4][G_M * G_ROW];

// k_gemm_outer with its stages (and, during the last stage, the C tile)
// brought into LDS by direct-to-LDS loads: no staging registers, no stores
// to LDS, and one wait per stage.  Wave w stages limb w of both operands, four
// 16-row chunks each: lane i covers row i / 4 of a chunk and LDS column i % 4,
// which holds source column (i % 4) ^ ((i / 16) % 4) (the swizzle of lds_at).
// The C tile (64 rows x 256 bytes) goes into the stage buffer the last stage
// does not read.  Requires 16-byte aligned C rows (the host checks; rows and
// columns past the matrix read clamped addresses and are never stored) and
// operand offsets below 2^32 bytes.
template <int JOB>
__global__ __launch_bounds__(G_T, 2) void k_gemm_outer_lds(const BMat<LimbsU32>* __restrict__ D, TileSched ts) {
    __shared__ __attribute__((aligned(16))) StageBuf sA0, sA1, sB0, sB1;
    int mi_;
    int64_t tm, tn;
    if (!batch_tile(ts, &mi_, &tm, &tn)) return;
    const BMat<LimbsU32>& d = D[mi_];
    uint32_t* C = JOB == 0 ? d.A + d.T0 : d.W;
    const int64_t ldc = JOB == 0 ? d.ld : BO;
    const int64_t Mrows = d.N;
    const int64_t Ncols = JOB == 0 ? d.nT : d.kA;
    const int64_t ldL = d.Kq;
    const int8_t* L = d.planesL + (JOB == 0 ? 0 : 16 * (int64_t)d.KqA);
    const int8_t* B = JOB == 0 ? d.planesB : d.planesW;
    const int64_t planeB = JOB == 0 ? d.strideB : STRIDE_W;
    const int Kpad = JOB == 0 ? d.Kq : d.Kq - d.KqA;
    const double pd = d.pd, pinv = d.pinv;
    const int64_t m_base = d.r0 + tm * G_M, n_base = tn * G_N;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int r16 = lane & 15, g = lane >> 4;
    const v4i zero = {0, 0, 0, 0};
    const int lrow = lane >> 2, lqu = (lane & 3) ^ ((lane >> 4) & 3);
    const int wu = __builtin_amdgcn_readfirstlane(wid);  // (the wave's limb, in scalar registers)
    const int8_t* la = L + wu * d.strideL;
    const int8_t* lb = B + wu * planeB;
    uint32_t oa[4], ob[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        oa[c] = (uint32_t)ptile(min(m_base + 16 * c + lrow, Mrows - 1), lqu * 16, ldL);
        ob[c] = (uint32_t)ptile(min(n_base + 16 * c + lrow, Ncols - 1), lqu * 16, Kpad);
    }
    // LDS byte address of this wave's limb in a stage buffer
    auto lds_of = [&](StageBuf& b) -> uint32_t {
        return __builtin_amdgcn_readfirstlane((uint32_t)(size_t)(__attribute__((address_space(3))) int8_t*)&b[wu][0]);
    };
    const uint32_t lA0 = lds_of(sA0), lA1 = lds_of(sA1), lB0 = lds_of(sB0), lB1 = lds_of(sB1);
    auto issue = [&](int kk, uint32_t dA, uint32_t dB) {  // (a stage is 1024 bytes of the tiled planes)
        const int8_t* sa = la + kk * 1024;
        const int8_t* sb = lb + kk * 1024;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            dma16_so(oa[c], sa, dA + c * 1024);
            dma16_so(ob[c], sb, dB + c * 1024);
        }
    };
    // wave w, chunk c: rows 16 w + 4 c .. + 3; lane i: row + i / 16, columns 4 (i % 16) .. + 3
    auto issue_c = [&](StageBuf& dst) {
        int8_t* base = &dst[0][0];
        const int64_t cmax = (Ncols - 1) / 4 * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int rl = 16 * wid + 4 * c + (lane >> 4);
            const int64_t row = min(m_base + rl, Mrows - 1);
            const int64_t col = min(n_base + (lane & 15) * 4, cmax);
            dma16(C + row * ldc + col, base + (16 * wid + 4 * c) * 256);
        }
    };
    v4i acc[7][2][2];
#pragma unroll
    for (int q = 0; q < 7; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[q][i][j] = zero;
    auto mfmas = [&](StageBuf& cA, StageBuf& cB) {
        v4i bf[4][2];
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
                bf[l][ni] = *reinterpret_cast<const v4i*>(&cB[l][lds_at(wn * 32 + ni * 16 + r16, g)]);
#pragma unroll
        for (int l1 = 0; l1 < 4; ++l1) {
            v4i af[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
                af[mi] = *reinterpret_cast<const v4i*>(&cA[l1][lds_at(wm * 32 + mi * 16 + r16, g)]);
#pragma unroll
            for (int l2 = 0; l2 < 4; ++l2)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int ni = 0; ni < 2; ++ni)
                        acc[l1 + l2][mi][ni] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(af[mi], bf[l2][ni], acc[l1 + l2][mi][ni], 0, 0, 0);
        }
    };
    const int nk = Kpad / G_K;
    auto stage = [&](int kk, StageBuf& cA, StageBuf& cB, StageBuf& nA, uint32_t nlA, uint32_t nlB) {
        lds_dma_barrier();
        if (kk + 1 < nk) issue(kk + 1, nlA, nlB);
        else issue_c(nA);
        mfmas(cA, cB);
    };
    issue(0, lA0, lB0);
    for (int kk = 0; kk < nk; kk += 2) {
        stage(kk, sA0, sB0, sA1, lA1, lB1);
        if (kk + 1 < nk) stage(kk + 1, sA1, sB1, sA0, lA0, lB0);
    }
    lds_dma_barrier();
    const uint32_t* sc = reinterpret_cast<const uint32_t*>(nk & 1 ? &sA1[0][0] : &sA0[0][0]);
    // one address and bound per (mi, ni); the lane's four rows i step by ldc
    // (no 64-bit multiply per element)
    const int rows_left = (int)min<int64_t>(Mrows - m_base, G_M);  // rows of the tile inside the matrix
    const int cols_left = (int)min<int64_t>(Ncols - n_base, G_N);
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
            const int rl0 = wm * 32 + mi * 16 + 4 * g, cl = wn * 32 + ni * 16 + r16;
            uint32_t* cp = C + (m_base + rl0) * ldc + n_base + cl;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (rl0 + i < rows_left && cl < cols_left) {
                    int32_t sv[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) sv[q] = acc[q][mi][ni][i];
                    gst<uint32_t>(cp + i * ldc, fold_sub(sc[(rl0 + i) * 64 + cl], sv, pd, pinv));
                }
            }
        }
}

// The 32-bit trailing GEMM launch: direct-to-LDS staging when allowed (tuning
// knob "gemm_lds" and the part's alignment), else register-staged copies.
template <int JOB>
static void launch_outer32(unsigned grid, hipStream_t st, const BMat<LimbsU32>* desc, const TileSched& ts, bool lds_ok) {
    if (lds_ok && tuning().gemm_lds != 0) k_gemm_outer_lds<JOB><<<grid, G_T, 0, st>>>(desc, ts);
    else k_gemm_outer<JOB><<<grid, G_T, 0, st>>>(desc, ts);
}

// Inner-step GEMM on the panel buffer:  X[r, :] -= X[r, c0 + Cp] U  for rows
// that are not pivots of this step (k_make_u has set X[R[t], :] = U[t, :] for
// those that are).  K = 64 (k <= 32 pivots, zero padded).  One workgroup owns a 64-row
// stripe and walks the live column window [c0, BO + q0) in 64-column tiles
// (columns outside it only meet zeros).  Owning the whole stripe lets it
// capture its own left operand, the digits of X[r, c0 + Cp], before any of its
// waves writes, and no other workgroup reads its rows.  Fragments come
// straight from global memory (LDS holds only the pivot columns): the launch is one workgroup per
// stripe, which matters because the trailing GEMM on the other stream holds
// every register of a CU and each workgroup here waits for one of its slots
// to free.  With few stripes (small matrices) `split` workgroups share a
// stripe, taking every split-th tile, and read the left operand from planes.  A tile's panel entries are in flight during its MFMAs, the next
// tile's right operand during its epilogue.
__global__ __launch_bounds__(G_T, 2) void k_gemm_inner(const BMat<LimbsU32>* __restrict__ D, TileSched ts, int epoch,
                                                     int split) {
    __shared__ int s_cp[CI];
    int mi_;
    int64_t tm, tn;
    if (!batch_tile(ts, &mi_, &tm, &tn)) return;
    const BMat<LimbsU32>& d = D[mi_];
    const BlockState<uint32_t>* bs = d.bs;
    const int k = bs->k;
    if (k == 0) return;
    const int c0 = bs->c0_cur;
    const int t_lo = c0 / G_N + (int)tn, t_hi = (BO + bs->q0 + G_N - 1) / G_N;
    if (t_lo >= t_hi) return;
    // addresses: a uniform stripe base plus 32-bit lane offsets
    const int64_t m_base = tm * G_M;
    const int rows = (int)min((int64_t)G_M, d.N - m_base);
    uint32_t* Xs = d.X + m_base * XLD;
    const int8_t* Ub = d.planesUi;
    const uint32_t p = d.f.p;
    const double pd = d.pd, pinv = d.pinv;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int r16 = lane & 15, g = lane >> 4;
    if (tid < CI) s_cp[tid] = tid < k ? bs->Cp[tid] : 0;
    // the stripe's inner panel columns [c0, c0 + BI), fetched row-contiguous
    // (rows past the matrix repeat its last row; their results are not stored)
    __shared__ uint32_t s_win[G_M][BI + 1];
    if (split == 1) {
        const int wr = tid >> 2, wq = tid & 3;  // row, eight-column quarter
        const uint32_t off = (uint32_t)min(wr, rows - 1) * XLD + c0 + wq * 8;
        const v4i w0 = gld<v4i>(Xs + off), w1 = gld<v4i>(Xs + off + 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            s_win[wr][wq * 8 + e] = (uint32_t)w0[e];
            s_win[wr][wq * 8 + 4 + e] = (uint32_t)w1[e];
        }
    }
    __syncthreads();
    const v4i zero = {0, 0, 0, 0};
    // left operand: byte j of af[l][mi] = digit l of X[row, c0 + Cp[t]], t = 16 g + j
    // (zero for t >= k).  With a split stripe another workgroup may already be
    // writing those columns, so the digits come from the planes k_make_u
    // captured before the launch.
    v4i af[4][2];
    if (split > 1) {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            const int64_t row = m_base + min(wm * 32 + mi * 16 + r16, rows - 1);
#pragma unroll
            for (int l = 0; l < 4; ++l) af[l][mi] = gld<v4i>(d.planesAi + l * d.strideAi + row * KPI + g * 16);
        }
    } else {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            const int rl = wm * 32 + mi * 16 + r16;
            uint32_t v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = s_win[rl][s_cp[(16 * g + j) & (CI - 1)]];
#pragma unroll
            for (int l = 0; l < 4; ++l) af[l][mi] = zero;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                int8_t a[4];
                balanced_digits(16 * g + j < k ? centred(v[j], p) : 0, a);
#pragma unroll
                for (int l = 0; l < 4; ++l) af[l][mi][j >> 2] |= (int32_t)(uint8_t)a[l] << (8 * (j & 3));
            }
        }
    }
    // per output row of this lane: its offset in the stripe, and whether it is
    // left alone (beyond the matrix, or a pivot row of this step: k_make_u has
    // written U into those)
    uint32_t ro[2][4];
    unsigned skip = 0;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int rl = wm * 32 + mi * 16 + 4 * g + i, rc = min(rl, rows - 1);
            ro[mi][i] = (uint32_t)rc * XLD + wn * 32 + r16;
            const bool piv = gld<int32_t>(d.epoch_of + (m_base + rc)) == epoch;
            skip |= (unsigned)(rl >= rows || piv) << (4 * mi + i);
        }
    const uint32_t bo = (uint32_t)(wn * 32 + r16) * KPI + g * 16;
    v4i bf[4][2];
    uint32_t cv[2][2][4];
    auto bload = [&](int t) {
#pragma unroll
        for (int l = 0; l < 4; ++l)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
                bf[l][ni] = gld<v4i>(Ub + (bo + (uint32_t)(l * XLD * KPI + (t * G_N + ni * 16) * KPI)));
    };
    bload(t_lo);
    __syncthreads();  // every wave has its left operand before any wave writes the stripe
    for (int t = t_lo; t < t_hi; t += split) {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)  // in flight during the MFMAs
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int i = 0; i < 4; ++i) cv[mi][ni][i] = gld<uint32_t>(Xs + (ro[mi][i] + t * G_N + ni * 16));
        __builtin_amdgcn_sched_barrier(0);  // keep those loads ahead of the MFMAs
        v4i acc[7][2][2];
#pragma unroll
        for (int s = 0; s < 7; ++s)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[s][i][j] = zero;
#pragma unroll
        for (int l1 = 0; l1 < 4; ++l1)
#pragma unroll
            for (int l2 = 0; l2 < 4; ++l2)
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int ni = 0; ni < 2; ++ni)
                        acc[l1 + l2][mi][ni] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(af[l1][mi], bf[l2][ni], acc[l1 + l2][mi][ni], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        bload(min(t + split, t_hi - 1));  // the next tile's right operand, in flight during the epilogue
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if ((skip >> (4 * mi + i)) & 1u) continue;
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
                    int32_t sv[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) sv[q] = acc[q][mi][ni][i];
                    gst<uint32_t>(Xs + (ro[mi][i] + t * G_N + ni * 16), fold_sub_k32(cv[mi][ni][i], sv, pd, pinv));
                }
            }
    }
}

// ---------------------------------------------------------------------------
// 64-bit storage: eight limb planes per operand and 15 diagonals.  A tile is
// 64 x TN64 (4 waves of 32 x 16 NI64); a 64-deep K stage is 8 x 8 x 2 NI64
// MFMAs per wave, long enough to hide the next stage's global loads behind a
// single LDS buffer (8 limbs x (64 + TN64) rows x 64 bytes).  With NI64 = 1
// the accumulators take 120 registers, so two workgroups share a CU and one's
// exact fold (about 150 VALU instructions per entry, mostly 64-bit integer
// work and int64 <-> double conversions) overlaps the other's MFMAs.
// ---------------------------------------------------------------------------

constexpr int NL64 = LimbsU64::NL, ND64 = 2 * NL64 - 1;
constexpr int NI64 = 1;             // 16-column blocks per wave
constexpr int TN64 = 32 * NI64;     // tile width
constexpr int BQ64 = TN64 / 8;      // B-stage 16-byte chunks per thread (TN64 rows x 4 x 8 limbs / 256)
constexpr int GW64 = NI64 == 1 ? 2 : 1;  // workgroups per CU (accumulators: 120 NI64 registers)

// acc[l1 + l2] += A_l1 B_l2 over one staged K slice (all 64 limb pairs); NI
// 16-column blocks per wave (tile width 32 NI).
template <int NI>
__device__ __forceinline__ void mfma_stage64(const int8_t (*sA)[G_M * G_ROW], const int8_t (*sB)[32 * NI * G_ROW],
                                             int wm, int wn, int r16, int g, v4i (&acc)[ND64][2][NI]) {
    v4i bf[NL64][NI];
#pragma unroll
    for (int l = 0; l < NL64; ++l)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
            bf[l][ni] = *reinterpret_cast<const v4i*>(&sB[l][lds_at(wn * 16 * NI + ni * 16 + r16, g)]);
#pragma unroll
    for (int l1 = 0; l1 < NL64; ++l1) {
        v4i af[2];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
            af[mi] = *reinterpret_cast<const v4i*>(&sA[l1][lds_at(wm * 32 + mi * 16 + r16, g)]);
#pragma unroll
        for (int l2 = 0; l2 < NL64; ++l2)
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
                    acc[l1 + l2][mi][ni] =
                        __builtin_amdgcn_mfma_i32_16x16x64_i8(af[mi], bf[l2][ni], acc[l1 + l2][mi][ni], 0, 0, 0);
    }
}

// B-stage chunks of a thread: chunk q is limb BL0 + q * BLS of row BROW,
// 16-byte column BQU (the same row and column for every q).
constexpr int BLS64 = G_T / (TN64 * 4);
static_assert(G_T % (TN64 * 4) == 0, "whole limbs per B-stage pass");

// Trailing-update GEMMs for 64-bit storage (JOB as in k_gemm_outer), tiles of
// 64 x 32 NI: NI = 1 (120 accumulator registers, two workgroups per CU) or
// NI = 2 (240, one workgroup per CU, half the right-operand stage loads per
// MAC); the tuning knob tile64 picks one for the trailing parts.
template <int NI>
struct Tile64 {
    static constexpr int TN = 32 * NI;           // tile width
    static constexpr int BQ = TN / 8;            // B-stage 16-byte chunks per thread (TN rows x 4 x 8 limbs / 256)
    static constexpr int BLS = G_T / (TN * 4);   // limb stride of a thread's B-stage chunks
    static constexpr int GW = NI == 1 ? 2 : 1;   // workgroups per CU
    static_assert(G_T % (TN * 4) == 0, "whole limbs per B-stage pass");
};

template <int JOB, int NI = 1>
__global__ __launch_bounds__(G_T, Tile64<NI>::GW) void k_gemm_outer64(const BMat<LimbsU64>* __restrict__ D,
                                                                     TileSched ts) {
    constexpr int TN = Tile64<NI>::TN, BQ = Tile64<NI>::BQ, BLS = Tile64<NI>::BLS;
    __shared__ __attribute__((aligned(16))) int8_t sA[NL64][G_M * G_ROW];
    __shared__ __attribute__((aligned(16))) int8_t sB[NL64][TN * G_ROW];
    int mi_;
    int64_t tm, tn;
    if (!batch_tile(ts, &mi_, &tm, &tn)) return;
    const BMat<LimbsU64>& d = D[mi_];
    uint64_t* C = JOB == 0 ? d.A + d.T0 : d.W;
    const int64_t ldc = JOB == 0 ? d.ld : BO;
    const int64_t Mrows = d.N;
    const int64_t Ncols = JOB == 0 ? d.nT : d.kA;
    const int64_t ldL = d.Kq;
    const int8_t* L = d.planesL + (JOB == 0 ? 0 : 16 * (int64_t)d.KqA);
    const int64_t planeL = d.strideL;
    const int8_t* B = JOB == 0 ? d.planesB : d.planesW;
    const int64_t planeB = JOB == 0 ? d.strideB : STRIDE_W;
    const int Kpad = JOB == 0 ? d.Kq : d.Kq - d.KqA;
    const uint64_t p = d.f.p;
    const double pinv = d.pinv;
    const int64_t m_base = d.r0 + tm * G_M, n_base = tn * TN;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int r16 = lane & 15, g = lane >> 4;
    const v4i zero = {0, 0, 0, 0};

    // A stage: limb q of row tid/4, quarter tid%4; B stage: BQ chunks
    v4i ra[NL64], rb[BQ];
    const int srow = tid >> 2, squ = tid & 3;
    const int8_t* pa = L + ptile(min(m_base + srow, Mrows - 1), squ * 16, ldL);
    const int soff = lds_at(srow, squ);
    const int bl0 = tid / (TN * 4), brow = (tid % (TN * 4)) >> 2, bqu = tid & 3;
    const int8_t* pb = B + bl0 * planeB + ptile(min(n_base + brow, Ncols - 1), bqu * 16, Kpad);
    const int boff = lds_at(brow, bqu);
    auto gload = [&](int kc) {  // (depth kc: 16 kc bytes into the tiled planes)
#pragma unroll
        for (int q = 0; q < NL64; ++q) ra[q] = gld<v4i>(pa + q * planeL + 16 * kc);
#pragma unroll
        for (int q = 0; q < BQ; ++q) rb[q] = gld<v4i>(pb + q * BLS * planeB + 16 * kc);
    };

    v4i acc[ND64][2][NI];
#pragma unroll
    for (int s = 0; s < ND64; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < NI; ++j) acc[s][i][j] = zero;

    const int nk = Kpad / G_K;
    gload(0);
    for (int kk = 0; kk < nk; ++kk) {
#pragma unroll
        for (int q = 0; q < NL64; ++q) *reinterpret_cast<v4i*>(&sA[q][soff]) = ra[q];
#pragma unroll
        for (int q = 0; q < BQ; ++q) *reinterpret_cast<v4i*>(&sB[bl0 + q * BLS][boff]) = rb[q];
        __syncthreads();
        if (kk + 1 < nk) gload((kk + 1) * G_K);
        mfma_stage64<NI>(sA, sB, wm, wn, r16, g, acc);
        __syncthreads();  // the next stage overwrites the buffer
    }
    // C values: issued together after the last stage (prefetching them into
    // the stage registers would keep 16 more addresses live through the loop,
    // beyond the register budget of two workgroups per CU); the other
    // workgroup's MFMAs cover the latency
    // (one address per 16 x 16 block, the lane's four rows ldc apart; rows
    // past the matrix read its last row)
    uint64_t cv[2][NI][4];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {  // unconditional loads, as in k_gemm_outer
            const int64_t row0 = m_base + wm * 32 + mi * 16 + 4 * g;
            const int64_t col = min(n_base + wn * 16 * NI + ni * 16 + r16, Ncols - 1);
            const uint64_t* cp = C + min(row0, Mrows - 1) * ldc + col;
            const uint64_t* last = C + (Mrows - 1) * ldc + col;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint64_t* q = cp + i * ldc;
                cv[mi][ni][i] = gld<uint64_t>(q < last ? q : last);
            }
        }
    const int rows_left = (int)min<int64_t>(Mrows - m_base, G_M);  // rows of the tile inside the matrix
    const int cols_left = (int)min<int64_t>(Ncols - n_base, 32 * NI);
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) {
            const int rl0 = wm * 32 + mi * 16 + 4 * g, cl = wn * 16 * NI + ni * 16 + r16;
            uint64_t* cp = C + (m_base + rl0) * ldc + n_base + cl;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (rl0 + i < rows_left && cl < cols_left) {
                    int32_t sv[ND64];
#pragma unroll
                    for (int q = 0; q < ND64; ++q) sv[q] = acc[q][mi][ni][i];
                    cp[i * ldc] = fold_sub64(cv[mi][ni][i], sv, p, pinv);
                }
            }
        }
}

// The 64-bit trailing GEMM with the tile width NI picked for the launch.
template <int JOB>
static void launch_outer64(unsigned grid, hipStream_t st, const BMat<LimbsU64>* desc, const TileSched& ts, int ni) {
    if (ni == 2) k_gemm_outer64<JOB, 2><<<grid, G_T, 0, st>>>(desc, ts);
    else k_gemm_outer64<JOB, 1><<<grid, G_T, 0, st>>>(desc, ts);
}

// Inner-step GEMM for 64-bit storage: as k_gemm_inner, a 64-row stripe per
// workgroup (or `split` workgroups per stripe, each taking every split-th
// tile) walking the live window in 64 x TN64 tiles.  The left operand comes
// from the planes k_make_u captured and is staged in LDS once; each tile
// stages its right operand.  Pivot rows of the step already hold U (k_make_u)
// and are skipped; so are columns past the window.
__global__ __launch_bounds__(G_T, GW64) void k_gemm_inner64(const BMat<LimbsU64>* __restrict__ D, TileSched ts,
                                                         int epoch, int split) {
    __shared__ __attribute__((aligned(16))) int8_t sA[NL64][G_M * G_ROW];
    __shared__ __attribute__((aligned(16))) int8_t sB[NL64][TN64 * G_ROW];
    int mi_;
    int64_t tm, tn;
    if (!batch_tile(ts, &mi_, &tm, &tn)) return;
    const BMat<LimbsU64>& d = D[mi_];
    const BlockState<uint64_t>* bs = d.bs;
    const int k = bs->k;
    if (k == 0) return;
    const int c0 = bs->c0_cur, colmax = BO + bs->q0;
    const int t_lo = c0 / TN64, t_hi = (colmax + TN64 - 1) / TN64;
    if (t_lo + (int)tn >= t_hi) return;
    const int64_t m_base = tm * G_M;
    uint64_t* X = d.X;
    const int64_t N = d.N;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int q = 0; q < NL64; ++q) {  // A: limb q, row tid / 4
        const int row_l = tid >> 2, qu = tid & 3;
        const int64_t row = min(m_base + row_l, N - 1);  // rows beyond N: clamped, results not stored
        *reinterpret_cast<v4i*>(&sA[q][lds_at(row_l, qu)]) = gld<v4i>(d.planesAi + q * d.strideAi + row * KPI + qu * 16);
    }
    // this lane's output rows that are left alone (beyond the matrix or pivot rows of the step)
    unsigned skip = 0;
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t row = m_base + wm * 32 + mi * 16 + 4 * g + i;
            const bool piv = gld<int32_t>(d.epoch_of + min(row, N - 1)) == epoch;
            skip |= (unsigned)(row >= N || piv) << (4 * mi + i);
        }
    const uint64_t p = d.f.p;
    const double pinv = d.pinv;
    const v4i zero = {0, 0, 0, 0};
    for (int t = t_lo + (int)tn; t < t_hi; t += split) {
        const int64_t n_base = (int64_t)t * TN64;
        __syncthreads();  // every wave is done with the previous tile's right operand
        {
            const int bl0 = tid / (TN64 * 4), bcol = (tid % (TN64 * 4)) >> 2, bqu = tid & 3;
#pragma unroll
            for (int q = 0; q < BQ64; ++q) {
                const int l = bl0 + q * BLS64;
                *reinterpret_cast<v4i*>(&sB[l][lds_at(bcol, bqu)]) =
                    gld<v4i>(d.planesUi + l * (int64_t)XLD * KPI + (n_base + bcol) * KPI + bqu * 16);
            }
        }
        // panel entries, in flight during the MFMAs
        uint64_t cv[2][NI64][4];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t row = min(m_base + wm * 32 + mi * 16 + 4 * g + i, N - 1);
#pragma unroll
                for (int ni = 0; ni < NI64; ++ni)
                    cv[mi][ni][i] = gld<uint64_t>(X + row * XLD + n_base + wn * 16 * NI64 + ni * 16 + r16);
            }
        __syncthreads();
        v4i acc[ND64][2][NI64];
#pragma unroll
        for (int s = 0; s < ND64; ++s)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < NI64; ++j) acc[s][i][j] = zero;
        mfma_stage64<NI64>(sA, sB, wm, wn, r16, g, acc);
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if ((skip >> (4 * mi + i)) & 1u) continue;
                const int64_t row = m_base + wm * 32 + mi * 16 + 4 * g + i;
#pragma unroll
                for (int ni = 0; ni < NI64; ++ni) {
                    const int64_t col = n_base + wn * 16 * NI64 + ni * 16 + r16;
                    if (col >= colmax) continue;
                    int32_t sv[ND64];
#pragma unroll
                    for (int q = 0; q < ND64; ++q) sv[q] = acc[q][mi][ni][i];
                    gst<uint64_t>(X + row * XLD + col, fold_sub64(cv[mi][ni][i], sv, p, pinv));
                }
            }
    }
}

// ---------------------------------------------------------------------------
// per-matrix set-up and status
// ---------------------------------------------------------------------------

// used = 0, epoch_of = -1, state reset, entries reduced mod p (blockIdx.y = matrix)
template <class LT>
__global__ __launch_bounds__(256) void k_item_init(const BMat<LT>* __restrict__ D) {
    const BMat<LT>& d = D[blockIdx.y];
    const int64_t g0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, gs = (int64_t)gridDim.x * blockDim.x;
    if (g0 == 0) {
        d.st->prow = -1;
        d.st->rank = 0;
        d.st->first_unused = 0;
        d.st->factor = 0;
    }
    for (int64_t r = g0; r < d.N; r += gs) {
        d.used[r] = 0;
        d.epoch_of[r] = -1;
    }
    // entries reduced mod p: a row per workgroup pass, four loads of a thread
    // in flight, written back only where they change (no division per entry;
    // loader output and canonical inputs are only read)
    using T = typename LT::T;
    for (int64_t r = blockIdx.x; r < d.N; r += gridDim.x) {
        T* row = d.A + r * d.ld;
        for (int64_t c0 = threadIdx.x; c0 < d.M; c0 += 4 * 256) {
            T v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = c0 + 256 * u < d.M ? row[c0 + 256 * u] : (T)0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const T y = d.f.reduce(v[u]);
                if (c0 + 256 * u < d.M && y != v[u]) row[c0 + 256 * u] = y;
            }
        }
    }
}

struct ItemStatus {
    int32_t c0_next, bcur, q0, pad;
    int64_t rank;
    int64_t lo;  // first unused row when the outer panel was loaded
};

// Ranks, then pivot lists, of many finished matrices in one launch each
// (blockIdx.x = matrix); `offs` are the prefix sums of the ranks.
template <class LT>
__global__ void k_collect_rank(const BMat<LT>* __restrict__ D, int64_t* out) {
    if (threadIdx.x == 0) out[blockIdx.x] = D[blockIdx.x].st->rank;
}

template <class LT>
__global__ __launch_bounds__(256) void k_collect_pivots(const BMat<LT>* __restrict__ D, const int64_t* __restrict__ offs,
                                                        int64_t* __restrict__ rows, int64_t* __restrict__ cols) {
    const BMat<LT>& d = D[blockIdx.x];
    const int64_t o = offs[blockIdx.x], n = offs[blockIdx.x + 1] - o;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
        rows[o + k] = d.piv_rows[k];
        cols[o + k] = d.piv_cols[k];
    }
}

// (Dl: the descriptors whose panel-load snapshot gives `lo`; null: D's own)
template <class LT>
__global__ void k_gather_status(const BMat<LT>* __restrict__ D, ItemStatus* out,
                                const BMat<LT>* __restrict__ Dl = nullptr) {
    const BMat<LT>& d = D[blockIdx.x];
    ItemStatus s;
    s.c0_next = d.bs->c0_next;
    s.bcur = d.bs->bcur;
    s.q0 = d.bs->q0;
    s.pad = 0;
    s.rank = d.st->rank;
    s.lo = (Dl ? Dl[blockIdx.x] : d).bs->lo0;
    out[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// Full-rank panel fast path.  If the first bcur unused rows (the candidate
// block) have full rank on the panel's bcur columns, every panel column is a
// pivot column, those rows can be the panel's pivot rows, and the panel's
// transformation restricted to them, Z[Q] = M^-1 with M = P[Q, pivot
// columns in slot order], follows from the candidate block alone.  Every other
// row r then has Z[r] = -P[r, :] B with B[c, s] = M^-1[slot of column c, s]
// (one GEMM) and a zero panel; the pivot rows get the block's rows.  The block
// is factorised by the inner-step kernels themselves, run on a compact copy
// (descriptors Dc, N = bcur); the regular inner steps over all N rows run only
// for matrices whose block is rank-deficient.  Over a field the RREF does not
// depend on which rows serve as pivots, so the result is the same.
// ---------------------------------------------------------------------------

// Candidate block of each matrix with Dc[m].N > 0: the first Dc[m].N unused
// rows (row order) into Dc[m].cand; the compact view's state reset (a matrix
// with Dc[m].N == 0 gets a finished block state, so that the compact inner
// steps skip it).  One workgroup of any whole number of waves (<= 16) per
// matrix, probing blockDim.x rows at a time; it runs as the last workgroup of
// the panel load (it reads only row flags and the first unused row, which the
// load does not touch).
template <class LT>
__device__ static void cand_gather_body(const BMat<LT>& d, const BMat<LT>& c) {
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
    if (tid == 0) {
        c.st->prow = -1;
        c.st->rank = 0;
        c.st->first_unused = 0;
        c.st->factor = 0;
        block_state_init<LT>(c, 0);
        if (c.N == 0) c.bs->c0_next = c.bcur;
    }
    if (c.N == 0) return;
    const int C = (int)c.N;
    __shared__ int s_total;
    __shared__ int s_wcnt[16];
    if (tid == 0) s_total = 0;
    __syncthreads();
    for (int64_t base = d.st->first_unused; base < d.N; base += blockDim.x) {
        const int64_t r = base + tid;
        const bool fl = r < d.N && !d.used[r];
        const unsigned long long m = __ballot(fl);
        if (lane == 0) s_wcnt[wid] = __popcll(m);
        __syncthreads();
        int before = s_total;
        for (int q = 0; q < wid; ++q) before += s_wcnt[q];
        const int pos = before + __popcll(m & ((1ull << lane) - 1ull));
        if (fl && pos < C) c.cand[pos] = r;
        __syncthreads();
        if (tid == 0) {
            int sum = 0;
            for (int q = 0; q < nw; ++q) sum += s_wcnt[q];
            s_total += sum;
        }
        __syncthreads();
        if (s_total >= C) break;
    }
    // (the host starts the fast path only where at least C rows are unused)
}

// The candidate rows' panel-buffer rows into the compact buffer, two rows
// per workgroup (blockIdx.x), and their compact row state reset.
template <class LT>
__global__ __launch_bounds__(256) void k_cand_copy(const BMat<LT>* __restrict__ D, const BMat<LT>* __restrict__ Dc) {
    using T = typename LT::T;
    const BMat<LT>& d = D[blockIdx.y];
    const BMat<LT>& c = Dc[blockIdx.y];
    constexpr int PER = 16 / (int)sizeof(T), VPR = XLD / PER;
    static_assert(2 * VPR <= 256 * 2, "two rows per workgroup");
    const int i = (int)blockIdx.x * 2 + (int)(threadIdx.x / 128);
    if (i >= c.N) return;
    const int64_t src = c.cand[i];
    for (int v = threadIdx.x % 128; v < VPR; v += 128)
        gst<v4i>(c.X + (int64_t)i * XLD + v * PER, gld<v4i>(d.X + src * XLD + v * PER));
    if (threadIdx.x % 128 == 0) {
        c.used[i] = 0;
        c.epoch_of[i] = -1;
        c.slot_of[i] = 0;
    }
}

// The candidate block's inner steps in one workgroup (32-bit storage), in
// place of k_inner_factor + k_make_u + k_gemm_inner per inner panel.  With a
// block of full rank the candidates of inner panel i are the block's rows
// [q0, q0 + BI) (every earlier group became pivots), so per inner panel:
//   1. fraction-free Gauss-Jordan on [X[cand, c0:c0+w] | I] in LDS (as in
//      k_inner_factor), normalisation, the pivots' bookkeeping and Z slots;
//   2. U = Apiv^-1 X[R, win] on the int8 matrix cores (limb planes of -Apiv^-1
//      and of X[R, win]^T in LDS, K = 32 slots), written to the pivot rows;
//   3. X[r, win] -= X[r, c0 + Cp] U for every other row of the block, a wave
//      per 32 rows (the left operand read before the wave writes its rows).
// win = the live columns [c0, BO + q0 + k) widened to multiples of 64.  A
// group that yields fewer pivots than its width stops the kernel; the block
// state then shows q0 < bcur and the host takes the regular path.
// Sized to share a CU with one workgroup of the trailing GEMM, which holds the
// rest of the device meanwhile: 8 waves of at most 128 registers, < 80 KB LDS.
// (Global writes of one wave are read by others only after a barrier: all
// waves of a workgroup share the CU's vector cache.)
constexpr int FB_WMAX = 320;  // widest window: BO + BI + 63 rounded to 64
constexpr int FB_WAVES = 8;

// Byte offset of the 8-byte K group g (slots 8 g .. 8 g + 7) of row `row` in a
// limb plane of 32-byte rows; the XOR keeps 16 consecutive rows of one group
// on distinct banks.
__device__ __forceinline__ int lds8(int row, int g) { return row * 32 + ((g ^ ((row >> 3) & 3)) << 3); }

typedef long v1l;  // the 8 int8 operand bytes of a 16x16x32 MFMA lane

__global__ __launch_bounds__(64 * FB_WAVES, 4) void k_block_factor(const BMat<LimbsU32>* __restrict__ Dc, int epoch0) {
    const BMat<LimbsU32>& c = Dc[blockIdx.x];
    if (c.N == 0) return;
    const FieldU32 f = c.f;
    const double pd = c.pd, pinv = c.pinv;
    uint32_t* __restrict__ X = c.X;
    const int C = (int)c.N, bcur = c.bcur;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int r16 = lane & 15, g = lane >> 4;
    constexpr int W2 = BI + CI, NW = FB_WAVES, RPW = CI / NW, NT = 64 * NW;
    static_assert(W2 == 64 && RPW * NW == CI, "one augmented column per lane, whole rows per wave");
    __shared__ uint32_t aug[CI][W2 + 1];
    __shared__ int prow_of[CI], pcol[CI], s_tof[CI];
    __shared__ uint32_t s_invm[CI];
    __shared__ __attribute__((aligned(16))) int8_t sA[4][CI * 32];       // -Apiv^-1: row t, K = slot
    __shared__ __attribute__((aligned(16))) int8_t sB[4][FB_WMAX * 32];  // X[R, win]^T, then U^T: row n, K = slot
    const v4i zero = {0, 0, 0, 0};
    int q0 = 0;
    int c0 = 0;
    unsigned long long tacc[6] = {0, 0, 0, 0, 0, 0};
    for (; c0 < bcur; c0 += BI) {
        const unsigned long long t0 = wall_clock64();
        const int w = min(BI, bcur - c0);
        const int nC = min(CI, C - q0);
        // 1. [X[q0 + i, c0:c0 + w] | I]; rows >= nC are zero and never pivot
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int ii = wid + NW * q, j = lane;
            uint32_t v = 0;
            if (ii < nC) {
                if (j < w) v = X[(int64_t)(q0 + ii) * XLD + c0 + j];
                else if (j >= BI) v = (j - BI == ii) ? 1u : 0u;
            }
            aug[ii][j] = v;
        }
        __syncthreads();
        unsigned long long pmask = 0ull;
        int t = 0;
        for (int cc = 0; cc < w && t < nC; ++cc) {
            uint32_t av[RPW], xv[RPW];
#pragma unroll
            for (int q = 0; q < RPW; ++q) {
                const int r = wid + NW * q;
                av[q] = aug[r][cc];
                xv[q] = aug[r][lane];
            }
            const bool nz = lane < nC && !((pmask >> lane) & 1ull) && aug[lane < CI ? lane : 0][cc] != 0u;
            const unsigned long long m = __ballot(nz);
            if (m == 0ull) continue;
            const int pv = __ffsll((long long)m) - 1;
            const uint32_t h = aug[pv][cc];
            const uint32_t y = aug[pv][lane];
#pragma unroll
            for (int q = 0; q < RPW; ++q) {
                const int r = wid + NW * q;
                if (r < nC && r != pv) aug[r][lane] = f.ff_update(h, xv[q], av[q], y);
            }
            if (tid == 0) {
                prow_of[t] = pv;
                pcol[t] = cc;
            }
            pmask |= 1ull << pv;
            ++t;
            __syncthreads();
        }
        const int k = t;
        __syncthreads();
        const unsigned long long t1 = wall_clock64();
        if (k < w) break;  // rank-deficient group: the host falls back (uniform: every wave found the same k)
        if (wid == 0 && lane < k) {
            const int r = prow_of[lane];
            s_invm[lane] = f.to_mont(f.inverse(aug[r][pcol[lane]]));
            s_tof[r] = lane;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPW; ++q) {
            const int r = wid + NW * q;
            if ((pmask >> r) & 1ull) aug[r][lane] = f.mont(s_invm[s_tof[r]], aug[r][lane]);
        }
        __syncthreads();
        // bookkeeping of the group's pivots; their Z slots become 1
        if (tid < k) {
            const int r = q0 + prow_of[tid];
            c.used[r] = 1;
            c.epoch_of[r] = epoch0 + c0 / BI;
            c.slot_of[r] = q0 + tid;
            c.piv_rows[q0 + tid] = r;
            c.piv_cols[q0 + tid] = c.J + c0 + pcol[tid];
            X[(int64_t)r * XLD + BO + q0 + tid] = 1u;
        }
        // limb planes of -Apiv^-1 (normal form): row t, K = slot s, 8 slots per store
        if (tid < CI * 4) {
            const int tt = tid >> 2, kg = tid & 3;
            int32_t pk[4][2] = {};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int sk = 8 * kg + j;
                const uint32_t v = (tt < k && sk < k) ? aug[prow_of[tt]][BI + prow_of[sk]] : 0u;
                int8_t d[4];
                balanced_digits(-centred(v, f.p), d);
#pragma unroll
                for (int l = 0; l < 4; ++l) pk[l][j >> 2] |= (int32_t)(uint8_t)d[l] << (8 * (j & 3));
            }
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                v1l w8;
                __builtin_memcpy(&w8, pk[l], 8);
                *reinterpret_cast<v1l*>(&sA[l][lds8(tt, kg)]) = w8;
            }
        }
        __syncthreads();  // (also orders the Z slot writes before the reads of X[R] below)
        const unsigned long long t2 = wall_clock64();
        const int qn = q0 + k;
        const int wlo = (c0 / 64) * 64, whi = (int)round_up(BO + qn, 64);
        const int W = whi - wlo;
        // limb planes of X[R, win]^T: row n, K = slot s; a thread per (n, 8-slot group)
        for (int x = tid; x < W * 4; x += NT) {
            const int n = x % W, kg = x / W;
            uint32_t v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int sk = 8 * kg + j;
                v[j] = X[(int64_t)(q0 + prow_of[sk < k ? sk : 0]) * XLD + wlo + n];
            }
            int32_t pk[4][2] = {};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                int8_t d[4];
                balanced_digits(8 * kg + j < k ? centred(v[j], f.p) : 0, d);
#pragma unroll
                for (int l = 0; l < 4; ++l) pk[l][j >> 2] |= (int32_t)(uint8_t)d[l] << (8 * (j & 3));
            }
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                v1l w8;
                __builtin_memcpy(&w8, pk[l], 8);
                *reinterpret_cast<v1l*>(&sB[l][lds8(n, kg)]) = w8;
            }
        }
        __syncthreads();
        const unsigned long long t3 = wall_clock64();
        // 2. U = Apiv^-1 X[R, win]: 16 x 16 output blocks (2 row halves x W / 16), spread over the waves
        const int nblk = 2 * (W / 16);
        for (int b = wid; b < nblk; b += NW) {
            const int th = b & 1, nb = b >> 1;
            v4i acc[7];
#pragma unroll
            for (int q = 0; q < 7; ++q) acc[q] = zero;
            v1l af[4], bf[4];
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                af[l] = *reinterpret_cast<const v1l*>(&sA[l][lds8(th * 16 + r16, g)]);
                bf[l] = *reinterpret_cast<const v1l*>(&sB[l][lds8(nb * 16 + r16, g)]);
            }
#pragma unroll
            for (int l1 = 0; l1 < 4; ++l1)
#pragma unroll
                for (int l2 = 0; l2 < 4; ++l2)
                    acc[l1 + l2] = __builtin_amdgcn_mfma_i32_16x16x32_i8(af[l1], bf[l2], acc[l1 + l2], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                int32_t sv[7];
#pragma unroll
                for (int q = 0; q < 7; ++q) sv[q] = acc[q][i];
                const int tt = th * 16 + 4 * g + i;  // 0 - sum(-Apiv^-1 X) = U, straight into the pivot row
                if (tt < k) X[(int64_t)(q0 + prow_of[tt]) * XLD + wlo + nb * 16 + r16] = fold_sub_k32(0u, sv, pd, pinv);
            }
        }
        __syncthreads();  // every wave is done reading sB
        const unsigned long long t4 = wall_clock64();
        // limb planes of U (now in the pivot rows; row n, K = slot t)
        for (int x = tid; x < W * 4; x += NT) {
            const int n = x % W, kg = x / W;
            uint32_t v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int tt = 8 * kg + j;
                v[j] = X[(int64_t)(q0 + prow_of[tt < k ? tt : 0]) * XLD + wlo + n];
            }
            int32_t pk[4][2] = {};
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                int8_t d[4];
                balanced_digits(8 * kg + j < k ? centred(v[j], f.p) : 0, d);
#pragma unroll
                for (int l = 0; l < 4; ++l) pk[l][j >> 2] |= (int32_t)(uint8_t)d[l] << (8 * (j & 3));
            }
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                v1l w8;
                __builtin_memcpy(&w8, pk[l], 8);
                *reinterpret_cast<v1l*>(&sB[l][lds8(n, kg)]) = w8;
            }
        }
        __syncthreads();
        const unsigned long long t5 = wall_clock64();
        // 3. the other rows of the block: wave wid owns rows 32 wid .. 32 wid + 31, in two groups of 16
#pragma unroll 1
        for (int grp = 0; grp < 2; ++grp) {
            const int rw = 32 * wid + 16 * grp;
            if (rw >= C || (rw < qn && rw + 16 > q0)) continue;  // beyond the block, or this group's pivots
            // left operand: byte j of af[l] = digit l of X[row, c0 + Cp[t]], t = 8 g + j
            v1l af[4];
            {
                const int row = min(rw + r16, C - 1);
                int32_t pk[4][2] = {};
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int tt = 8 * g + j;
                    const uint32_t v = tt < k ? X[(int64_t)row * XLD + c0 + pcol[tt]] : 0u;
                    int8_t d[4];
                    balanced_digits(centred(v, f.p), d);
#pragma unroll
                    for (int l = 0; l < 4; ++l) pk[l][j >> 2] |= (int32_t)(uint8_t)d[l] << (8 * (j & 3));
                }
#pragma unroll
                for (int l = 0; l < 4; ++l) __builtin_memcpy(&af[l], pk[l], 8);
            }
            uint32_t* xb = X + (int64_t)rw * XLD + wlo + r16;  // the group's rows: 32-bit offsets from here
            uint32_t xo[4], cn[4];  // cn: C values of the next block, in flight during this one
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                xo[i] = (uint32_t)(min(rw + 4 * g + i, C - 1) - rw) * XLD;
                cn[i] = xb[xo[i]];
            }
            for (int nb = 0; nb < W / 16; ++nb) {
                uint32_t cv[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    cv[i] = cn[i];
                    cn[i] = xb[xo[i] + (uint32_t)min(nb + 1, W / 16 - 1) * 16];
                }
                v1l bf[4];
#pragma unroll
                for (int l = 0; l < 4; ++l) bf[l] = *reinterpret_cast<const v1l*>(&sB[l][lds8(nb * 16 + r16, g)]);
                v4i acc[7];
#pragma unroll
                for (int q = 0; q < 7; ++q) acc[q] = zero;
#pragma unroll
                for (int l1 = 0; l1 < 4; ++l1)
#pragma unroll
                    for (int l2 = 0; l2 < 4; ++l2)
                        acc[l1 + l2] = __builtin_amdgcn_mfma_i32_16x16x32_i8(af[l1], bf[l2], acc[l1 + l2], 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (rw + 4 * g + i >= C) continue;
                    int32_t sv[7];
#pragma unroll
                    for (int q = 0; q < 7; ++q) sv[q] = acc[q][i];
                    xb[xo[i] + nb * 16] = fold_sub_k32(cv[i], sv, pd, pinv);
                }
            }
        }
        __syncthreads();
        const unsigned long long t6 = wall_clock64();
        tacc[0] += t1 - t0;
        tacc[1] += t2 - t1;
        tacc[2] += t3 - t2;
        tacc[3] += t4 - t3;
        tacc[4] += t5 - t4;
        tacc[5] += t6 - t5;
        q0 = qn;
    }
    if (tid == 0) {
        for (int i = 0; i < 6; ++i) c.bs->tphase[i] += tacc[i];
        c.bs->tphase[6] += 1;
        c.st->rank = q0;
        c.bs->q0 = q0;
        c.bs->k = 0;
        c.bs->c0_cur = c0 < bcur ? c0 : 0;
        c.bs->c0_next = c0 < bcur ? c0 : bcur;
    }
}

// The fast path's kernels are queued before the host knows which candidate
// blocks reached full rank: this gate gives every other matrix nT = 0 (the
// fast kernels then skip it) and no tiles in the Z GEMM's schedule
// (prefix sums pre[0..B] rewritten in place).  One thread.
// (and, in the same launch, the compact blocks' status with the main
// descriptors' panel-load snapshot of the first unused row, as k_gather_status)
template <class LT>
__global__ __launch_bounds__(256) void k_fast_gate(BMat<LT>* __restrict__ Dw, const BMat<LT>* __restrict__ Dc,
                                                   int64_t* __restrict__ pre, int B, ItemStatus* __restrict__ out,
                                                   const BMat<LT>* __restrict__ Dl) {
    for (int m = threadIdx.x; m < B; m += blockDim.x) {
        const BMat<LT>& c = Dc[m];
        ItemStatus st;
        st.c0_next = c.bs->c0_next;
        st.bcur = c.bs->bcur;
        st.q0 = c.bs->q0;
        st.pad = 0;
        st.rank = c.st->rank;
        st.lo = Dl[m].bs->lo0;
        out[m] = st;
    }
    // drop the blocks that missed full rank and rewrite the Z GEMM's tile
    // prefix sums in place over the kept ones: for a few matrices one thread
    // in turn; for more, 256 matrices at a time, their reads all in flight
    // together, an inclusive scan in LDS (the serial loop chains several
    // dependent loads per matrix)
    if (B <= 8) {
        if (threadIdx.x != 0) return;
        int64_t kept = 0, prev = pre[0];
        for (int m = 0; m < B; ++m) {
            const int64_t tiles = pre[m + 1] - prev;
            prev = pre[m + 1];
            const int bcur = Dw[m].bcur;
            const bool full = Dw[m].nT > 0 && Dc[m].st->rank == bcur && Dc[m].bs->q0 == bcur;
            if (!full) Dw[m].nT = 0;
            kept += full ? tiles : 0;
            pre[m + 1] = kept;
        }
        return;
    }
    __shared__ int64_t s_scan[256];
    __shared__ int64_t s_run, s_edge;  // kept tiles before this chunk; the chunk's last original prefix
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_run = 0;
        s_edge = pre[0];
    }
    __syncthreads();
    for (int base = 0; base < B; base += 256) {
        const int m = base + tid;
        int64_t t = 0, hi = 0;
        if (m < B) {
            hi = pre[m + 1];
            const int64_t lo = tid == 0 ? s_edge : pre[m];
            const int bcur = Dw[m].bcur;
            const bool full = Dw[m].nT > 0 && Dc[m].st->rank == bcur && Dc[m].bs->q0 == bcur;
            t = full ? hi - lo : 0;
            if (!full) Dw[m].nT = 0;
        }
        s_scan[tid] = t;
        __syncthreads();  // (every original prefix of the chunk is read before any is rewritten)
        if (tid == 255 || m == B - 1) s_edge = hi;
#pragma unroll
        for (int off = 1; off < 256; off <<= 1) {
            const int64_t o = tid >= off ? s_scan[tid - off] : 0;
            __syncthreads();
            s_scan[tid] += o;
            __syncthreads();
        }
        if (m < B) pre[m + 1] = s_run + s_scan[tid];
        __syncthreads();
        if (tid == 0) s_run += s_scan[255];
        __syncthreads();
    }
}

// Operands of Z = -P B for the matrices on the fast path (Dw[m].nT > 0, depth
// Kq = bcur rounded up to 64): workgroups [0, nl) write the limb planes of
// the panel P = X[:, 0:bcur] (every row from r0 on, zero beyond bcur) with row stride Kq
// into planesL1 and set those panel columns to zero; the rest write those of
// B^T, B[c, s] = M^-1[slot of column c, s] taken from the factorised compact
// block, into planesB1 (row s, stride Kq).
template <class LT>
__global__ __launch_bounds__(256) void k_fast_operands(const BMat<LT>* __restrict__ Dw, const BMat<LT>* __restrict__ Dc,
                                                       int nl) {
    constexpr int NL = LT::NL;
    const BMat<LT>& d = Dw[blockIdx.y];
    if (d.nT == 0) return;
    const int Kq = d.Kq, bcur = d.bcur;
    const int64_t bx = blockIdx.x;
    if (bx < nl) {
        const int64_t total = (d.N - d.r0) * (Kq / 4);
        for (int64_t x = bx * 256 + threadIdx.x; x < total; x += (int64_t)nl * 256) {
            const int64_t i = x / (Kq / 4), r = d.r0 + i;
            const int c = (int)(x - i * (Kq / 4)) * 4;
            int8_t dg[NL][4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                int8_t g[NL];
                balanced_digits(c + e < bcur ? centred(d.X[r * XLD + c + e], d.f.p) : (typename LT::S)0, g);
#pragma unroll
                for (int l = 0; l < NL; ++l) dg[l][e] = g[l];
            }
            // the panel columns of every such row end as zero (the pivot rows
            // get theirs back from the block in k_fast_commit); c < Kq <= BO
#pragma unroll
            for (int e = 0; e < 4; ++e) d.X[r * XLD + c + e] = (typename LT::T)0;
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                int32_t pk;
                __builtin_memcpy(&pk, dg[l], 4);
                *reinterpret_cast<int32_t*>(d.planesL1 + l * d.strideL1 + ptile(r, c, Kq)) = pk;
            }
        }
        return;
    }
    const BMat<LT>& c = Dc[blockIdx.y];
    const int64_t nb = gridDim.x - nl;
    for (int64_t x = (bx - nl) * 256 + threadIdx.x; x < (int64_t)bcur * bcur; x += nb * 256) {
        const int t = (int)(x / bcur), sidx = (int)(x - (int64_t)t * bcur);
        const int col = (int)(c.piv_cols[t] - c.J);
        int8_t g[NL];
        balanced_digits(centred(c.X[c.piv_rows[t] * XLD + BO + sidx], d.f.p), g);
#pragma unroll
        for (int l = 0; l < NL; ++l) d.planesB1[l * STRIDE_W + ptile(sidx, col, Kq)] = g[l];
    }
}

// The block's pivot rows back into the panel buffer (normalised panel row
// and M^-1 row) and their bookkeeping -- row flags, epochs, slots, pivot lists
// -- two pivots per workgroup (blockIdx.x); the last workgroup advances the
// first unused row and sets rank and block state.
template <class LT>
__global__ __launch_bounds__(256) void k_fast_commit(const BMat<LT>* __restrict__ D, const BMat<LT>* __restrict__ Dc,
                                                     const BMat<LT>* __restrict__ Dw, int epoch) {
    using T = typename LT::T;
    const BMat<LT>& d = D[blockIdx.y];
    const BMat<LT>& c = Dc[blockIdx.y];
    if (Dw[blockIdx.y].nT == 0) return;
    const int q = (int)c.st->rank;  // == bcur on the fast path
    if (blockIdx.x == gridDim.x - 1) {
        // the bookkeeping: rank, block state and the first unused row, which
        // lies beyond the block's last row (the block is the first q unused
        // rows, so every row up to its last one is used once it is committed)
        // -- found from rows no other workgroup of this launch writes
        const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
        __shared__ unsigned long long s_first;
        if (tid == 0) s_first = ~0ull;
        __syncthreads();
        for (int64_t base = c.cand[q - 1] + 1; base < d.N; base += blockDim.x) {
            const int64_t r = base + tid;
            const unsigned long long m = __ballot(r < d.N && !d.used[r]);
            if (lane == 0 && m) atomicMin(&s_first, (unsigned long long)(base + wid * 64 + __ffsll((long long)m) - 1));
            __syncthreads();
            const bool found = s_first != ~0ull;
            __syncthreads();
            if (found) break;
        }
        if (tid == 0) {
            d.st->rank = d.rank0 + q;
            d.st->first_unused = s_first != ~0ull ? (int64_t)s_first : d.N;
            d.bs->k = 0;
            d.bs->q0 = q;
            d.bs->c0_cur = 0;
            d.bs->c0_next = d.bcur;
        }
        return;
    }
    const int t = (int)blockIdx.x * 2 + (int)(threadIdx.x / 128);
    if (t >= q) return;
    constexpr int PER = 16 / (int)sizeof(T), VPR = XLD / PER;
    const int64_t cr = c.piv_rows[t];
    const int64_t gr = c.cand[cr];
    for (int v = threadIdx.x % 128; v < VPR; v += 128)
        gst<v4i>(d.X + gr * XLD + v * PER, gld<v4i>(c.X + cr * XLD + v * PER));
    if (threadIdx.x % 128 == 0) {
        const int64_t rank0 = d.rank0;
        d.used[gr] = 1;
        d.epoch_of[gr] = epoch;
        d.slot_of[gr] = t;
        d.piv_rows[rank0 + t] = gr;
        d.piv_cols[rank0 + t] = c.piv_cols[t];
    }
}

// ---------------------------------------------------------------------------
// host engine: a set of matrices advanced together, one outer panel per step
// ---------------------------------------------------------------------------

struct WorkspaceLayout {
    size_t X, L, Wp, B, L1, B1, Ai, Ui, U, W, bs, st, piv, used, epoch, slot;
    // the compact candidate block (BO rows) of the full-rank panel fast path
    size_t cX, cAi, cUi, cU, cbs, cst, cpiv, cused, cepoch, cslot, cand;
    size_t total;
};

template <class LT>
static WorkspaceLayout workspace_layout(int64_t N, int64_t M, bool with_elim_state) {
    using T = typename LT::T;
    constexpr size_t NL = LT::NL, E = sizeof(T);
    auto al = [](size_t x) { return (size_t)round_up((int64_t)x, 256); };
    WorkspaceLayout w{};
    const int64_t cap = std::max<int64_t>(1, std::min(N, M));
    size_t off = 0;
    w.X = off; off += al((size_t)N * XLD * E);
    w.L = off; off += al(NL * round_up(prows(N) * 2 * BO, 256));
    w.Wp = off; off += al(NL * STRIDE_W);
    w.B = off; off += al(NL * round_up(prows(std::max<int64_t>(M, XLD)) * 2 * BO, 256));
    w.L1 = off; off += al(NL * round_up(prows(N) * BO, 256));
    w.B1 = off; off += al(NL * STRIDE_W);
    w.Ai = off; off += al(NL * round_up(N * KPI, 256));
    w.Ui = off; off += al(NL * XLD * KPI);
    w.U = off; off += al((size_t)CI * XLD * E);
    w.W = off; off += al((size_t)N * BO * E);
    w.bs = off; off += al(sizeof(BlockState<T>));
    w.epoch = off; off += al((size_t)4 * N);
    w.slot = off; off += al((size_t)4 * N);
    if (with_elim_state) {
        w.st = off; off += al(sizeof(ElimState));
        w.piv = off; off += 2 * al((size_t)cap * 8);
        w.used = off; off += al((size_t)N);
    }
    w.cX = off; off += al((size_t)BO * XLD * E);
    w.cAi = off; off += al(NL * BO * KPI);
    w.cUi = off; off += al(NL * XLD * KPI);
    w.cU = off; off += al((size_t)CI * XLD * E);
    w.cbs = off; off += al(sizeof(BlockState<T>));
    w.cst = off; off += al(sizeof(ElimState));
    w.cpiv = off; off += 2 * al((size_t)BO * 8);
    w.cused = off; off += al(BO);
    w.cepoch = off; off += al((size_t)4 * BO);
    w.cslot = off; off += al((size_t)4 * BO);
    w.cand = off; off += al((size_t)8 * BO);
    w.total = off;
    return w;
}

template <class LT>
static void carve(BMat<LT>& d, char* base, const WorkspaceLayout& w, int64_t N, int64_t M, bool with_elim_state) {
    using T = typename LT::T;
    d.X = (T*)(base + w.X);
    d.planesL = (int8_t*)(base + w.L);
    d.strideL = round_up(prows(N) * 2 * BO, 256);
    d.planesW = (int8_t*)(base + w.Wp);
    d.planesB = (int8_t*)(base + w.B);
    d.strideB = round_up(prows(std::max<int64_t>(M, XLD)) * 2 * BO, 256);
    d.planesL1 = (int8_t*)(base + w.L1);
    d.strideL1 = round_up(prows(N) * BO, 256);
    d.planesB1 = (int8_t*)(base + w.B1);
    d.planesAi = (int8_t*)(base + w.Ai);
    d.strideAi = round_up(N * KPI, 256);
    d.planesUi = (int8_t*)(base + w.Ui);
    d.U = (T*)(base + w.U);
    d.W = (T*)(base + w.W);
    d.bs = (BlockState<T>*)(base + w.bs);
    d.epoch_of = (int32_t*)(base + w.epoch);
    d.slot_of = (int32_t*)(base + w.slot);
    if (with_elim_state) {
        const int64_t cap = std::max<int64_t>(1, std::min(N, M));
        d.st = (ElimState*)(base + w.st);
        d.piv_rows = (int64_t*)(base + w.piv);
        d.piv_cols = (int64_t*)(base + w.piv + round_up(cap * 8, 256));
        d.used = (uint8_t*)(base + w.used);
    }
    d.cand = (int64_t*)(base + w.cand);
}

// The compact view of the candidate block: the inner-step kernels' fields
// only (panel buffer, inner GEMM planes, U, state), BO rows.
template <class LT>
static void carve_compact(BMat<LT>& c, char* base, const WorkspaceLayout& w) {
    using T = typename LT::T;
    c = BMat<LT>{};
    c.X = (T*)(base + w.cX);
    c.planesAi = (int8_t*)(base + w.cAi);
    c.strideAi = round_up((int64_t)BO * KPI, 256);
    c.planesUi = (int8_t*)(base + w.cUi);
    c.U = (T*)(base + w.cU);
    c.bs = (BlockState<T>*)(base + w.cbs);
    c.st = (ElimState*)(base + w.cst);
    c.piv_rows = (int64_t*)(base + w.cpiv);
    c.piv_cols = (int64_t*)(base + w.cpiv + round_up((int64_t)BO * 8, 256));
    c.used = (uint8_t*)(base + w.cused);
    c.epoch_of = (int32_t*)(base + w.cepoch);
    c.slot_of = (int32_t*)(base + w.cslot);
    c.cand = (int64_t*)(base + w.cand);
}

template <class LT>
struct Item {
    BMat<LT> d{};
    BMat<LT> dc{};             // compact view of the candidate block (fast path)
    int64_t rank = 0;          // rank after the last step, as the host knows it
    void* arena = nullptr;     // device memory owned by the item (null when borrowed)
    size_t bytes = 0;
    bool finished = false;
    bool taken = false;
    int64_t id = -1;
    // Forward elimination: each panel's updates start at row r0 (a panel that
    // composes with a deferred one keeps that one's r0); the panels that
    // skipped rows go to `echelon` for the back substitution.
    bool forward = false;
    int64_t r0 = 0;
    std::vector<EchelonPanel> echelon;
};

template <class LT>
class Engine {
  public:
    using BM = BMat<LT>;
    using It = Item<LT>;
    Engine() = default;
    ~Engine() {
        if (so) (void)hipStreamSynchronize(so);
        for (int k = 0; k < NSLOT; ++k) {
            if (d_buf[k]) (void)hipFree(d_buf[k]);
            if (h_buf[k]) (void)hipHostFree(h_buf[k]);
        }
        if (ev_ready) (void)hipEventDestroy(ev_ready);
        if (ev_status) (void)hipEventDestroy(ev_status);
        if (ev_inner) (void)hipEventDestroy(ev_inner);
        for (int i = 0; i < NEV; ++i) {
            if (ev_f[i]) (void)hipEventDestroy(ev_f[i]);
            if (ev_c[i]) (void)hipEventDestroy(ev_c[i]);
        }
        if (so) (void)hipStreamDestroy(so);
    }
    // One outer panel for every unfinished item in `items`.  The inner phase
    // runs on `s`; the trailing update on the engine's own stream, where the
    // part beyond the next panel keeps running after step() returns.
    int step(const std::vector<It*>& items, hipStream_t s, Profiler* prof);
    int init_items(const std::vector<It*>& items, hipStream_t s);
    // Waits for every trailing update the engine has issued (before results
    // are read or an item's memory is released).
    int drain();
    int epoch = 0;
    bool first_step = true;
    // Ready for a new solve after drain(): the engine (its stream, events and
    // staging buffers) is cached per device and reused across calls.
    void reset_for_solve() {
        epoch = 0;
        first_step = true;
        step_no = 0;
        bulk.clear();
        ev_pending = false;
    }

  private:
    // Staging slots (pinned host + device): 0 = inner phase (with the fast
    // path's compact views and Z-GEMM descriptors); 1, 2 = the three parts of
    // the trailing update of even and odd steps.  Slot 0 is refilled on `s`
    // after every kernel of the previous step that read it.  A slot's device
    // copy is refilled on `so` after every kernel that read it, and its host
    // side two steps after its copy was queued (the inner phase of a step
    // waits for the previous step's first part, which follows the parts
    // before it on the same stream).
    static constexpr int NSLOT = 3;
    struct Staged {
        BM* desc;
        TileSched inner, wq, main;  // inner GEMM, composite W GEMM, trailing update
        int64_t inner_tiles, wq_tiles, main_tiles;
        int inner_split;            // 32-bit inner GEMM: workgroups per row stripe
        ItemStatus* status;         // device copy; status_h: its pinned host mirror
        ItemStatus* status_h;
        double macs, bytes;         // algorithmic work of the trailing update (profiler)
        int64_t cols;               // widest column range of the part
        bool lds_main, lds_wq;      // 32-bit trailing / composite GEMMs may stage through direct-to-LDS loads
        int ni64;                   // 64-bit trailing / composite GEMMs: tile width 32 ni64
    };
    int stage(int slot, size_t bytes);
    // part 0: inner phase; 1: the next panel's columns; 2, 3: the two halves
    // of the columns beyond it (composite updates only)
    // parts part .. part + nparts - 1 into one staging slot, one copy
    int upload(const std::vector<It*>& act, hipStream_t s, int slot, int part, Staged* out, int nparts = 1);
    // the listed parts (at most 8) into one staging slot, one copy
    int upload_list(const std::vector<It*>& act, hipStream_t s, int slot, const int* parts, int nparts, Staged* out);
    // parts 4 (compact block views) and 5 (fast-path GEMM Z = -P B): which items take part
    std::vector<char> sel;
    std::vector<int64_t> lo_fast;  // first unused row at the panel load, read with the fast path's status
    double spin_us = 0;            // host wait for the last status read (debug output)
    int fast_panel(const std::vector<It*>& act, hipStream_t s, const Staged& in, const Staged& cs, const Staged& ws,
                   int64_t ninner, int64_t maxN, std::vector<char>& fast, Profiler* prof);
    void fill(const std::vector<It*>& act, int part, char* h, char* dv, size_t desc_b, size_t sched_b, Staged* out);
    int trailing(const Staged& t, Profiler* prof, int64_t tag, bool rhs_ready = false);
    // A trailing part beyond the next panel, queued to run beside a later
    // step's inner phase (`due`), launched in tile slices.
    struct Bulk {
        Staged t;
        int64_t next;  // tiles launched so far
        int64_t due;
    };
    int launch_slice(Bulk& b, int64_t ntiles, Profiler* prof, int64_t tag);  // ntiles < 0: the rest
    int flush_due(int64_t upto, Profiler* prof, int64_t tag);                  // every part due by `upto`
    int ensure_stream();
    hipStream_t so = nullptr;
    hipEvent_t ev_ready = nullptr;  // on `so`: next panel's columns updated, panel buffer free
    hipEvent_t ev_status = nullptr;  // on `s`: the inner phase's status has reached host memory
    hipEvent_t ev_inner = nullptr;   // on `s`: the inner phase is complete (fast path without a status read)
    bool ev_pending = false;
    std::vector<Bulk> bulk;         // queued trailing parts, in due order
    int64_t step_no = 0;
    static constexpr int NEV = 16;  // slices per step (inner groups beyond this run unsliced)
    hipEvent_t ev_f[NEV] = {}, ev_c[NEV] = {};  // inner step started (on s) / its slice done (on so)
    int parity = 0;
    void* d_buf[NSLOT] = {};
    void* h_buf[NSLOT] = {};
    size_t cap[NSLOT] = {};
};

template <class LT>
int Engine<LT>::ensure_stream() {
    if (so) return LINAC_OK;
    LINAC_HIP(hipStreamCreateWithFlags(&so, hipStreamNonBlocking));
    LINAC_HIP(hipEventCreateWithFlags(&ev_ready, hipEventDisableTiming));
    LINAC_HIP(hipEventCreateWithFlags(&ev_status, hipEventDisableTiming));
    LINAC_HIP(hipEventCreateWithFlags(&ev_inner, hipEventDisableTiming));
    for (int i = 0; i < NEV; ++i) {
        LINAC_HIP(hipEventCreateWithFlags(&ev_f[i], hipEventDisableTiming));
        LINAC_HIP(hipEventCreateWithFlags(&ev_c[i], hipEventDisableTiming));
    }
    return LINAC_OK;
}

template <class LT>
int Engine<LT>::drain() {
    LINAC_TRY(flush_due(INT64_MAX, nullptr, 0));
    if (so) LINAC_HIP(hipStreamSynchronize(so));
    return LINAC_OK;
}

// Tiles [next, next + ntiles) of a queued part's trailing GEMM on `so` (its
// B operand planes before the first slice).
template <class LT>
int Engine<LT>::launch_slice(Bulk& b, int64_t ntiles, Profiler* prof, int64_t tag) {
    const Staged& t = b.t;
    const int64_t rest = t.main_tiles - b.next;
    if (rest <= 0) return LINAC_OK;
    const int64_t n = ntiles < 0 ? rest : std::min(rest, ntiles);
    if (b.next == 0)
        k_make_rhs<LT><<<dim3((unsigned)std::min<int64_t>(2048, ceil_div(t.cols, 64)), 2 * BO / 64,
                              (unsigned)t.inner.B), 256, 0, so>>>(t.desc, 0);
    TileSched ts = t.main;
    ts.t0 = b.next;
    ts.t1 = b.next + n;
    const double frac = (double)n / (double)t.main_tiles;
    if (prof) LINAC_TRY(prof->sample_begin(so, tag, t.bytes * frac, t.macs * frac));
    const unsigned grid = (unsigned)(ceil_div(n, NUM_XCD) * NUM_XCD);
    if constexpr (std::is_same<LT, LimbsU64>::value) launch_outer64<0>(grid, so, t.desc, ts, t.ni64);
    else launch_outer32<0>(grid, so, t.desc, ts, t.lds_main);
    if (prof) LINAC_TRY(prof->sample_end(so));
    LINAC_HIP(hipGetLastError());
    b.next += n;
    return LINAC_OK;
}

template <class LT>
int Engine<LT>::flush_due(int64_t upto, Profiler* prof, int64_t tag) {
    size_t done = 0;
    for (auto& b : bulk) {
        if (b.due > upto) break;
        LINAC_TRY(launch_slice(b, -1, prof, tag));
        ++done;
    }
    bulk.erase(bulk.begin(), bulk.begin() + done);
    return LINAC_OK;
}

// The trailing GEMM of one staged part on `so`: B operand planes, then the update.
template <class LT>
int Engine<LT>::trailing(const Staged& t, Profiler* prof, int64_t tag, bool rhs_ready) {
    if (t.main_tiles == 0) return LINAC_OK;
    const unsigned B = (unsigned)t.inner.B;
    if (!rhs_ready)
        k_make_rhs<LT><<<dim3((unsigned)std::min<int64_t>(2048, ceil_div(t.cols, 64)), 2 * BO / 64, B), 256, 0, so>>>(
            t.desc, 0);
    if (prof) LINAC_TRY(prof->sample_begin(so, tag, t.bytes, t.macs));
    const unsigned grid = (unsigned)(ceil_div(t.main_tiles, NUM_XCD) * NUM_XCD);
    if constexpr (std::is_same<LT, LimbsU64>::value) launch_outer64<0>(grid, so, t.desc, t.main, t.ni64);
    else launch_outer32<0>(grid, so, t.desc, t.main, t.lds_main);
    if (prof) LINAC_TRY(prof->sample_end(so));
    LINAC_HIP(hipGetLastError());
    return LINAC_OK;
}

template <class LT>
int Engine<LT>::stage(int slot, size_t bytes) {
    if (bytes <= cap[slot]) return LINAC_OK;
    if (so) LINAC_HIP(hipStreamSynchronize(so));
    if (d_buf[slot]) LINAC_HIP(hipFree(d_buf[slot]));
    if (h_buf[slot]) LINAC_HIP(hipHostFree(h_buf[slot]));
    d_buf[slot] = h_buf[slot] = nullptr;
    cap[slot] = (size_t)round_up((int64_t)bytes * 2, 4096);
    LINAC_HIP(hipMalloc(&d_buf[slot], cap[slot]));
    LINAC_HIP(hipHostMalloc(&h_buf[slot], cap[slot], hipHostMallocDefault));
    return LINAC_OK;
}

// Staging layout: [descriptors B][sched 6B+3 int64][status B].  The sched
// arrays are prefix sums of per-matrix tile counts: (pa, ma) inner GEMM,
// (pb, mb) composite W of a deferred panel, (pc, mc) trailing update.
template <class LT>
int Engine<LT>::upload(const std::vector<It*>& act, hipStream_t s, int slot, int part, Staged* out, int nparts) {
    int parts[8];
    for (int i = 0; i < nparts; ++i) parts[i] = part + i;
    return upload_list(act, s, slot, parts, nparts, out);
}

template <class LT>
int Engine<LT>::upload_list(const std::vector<It*>& act, hipStream_t s, int slot, const int* parts, int nparts,
                            Staged* out) {
    const size_t B = act.size();
    const size_t desc_b = (size_t)round_up((int64_t)(B * sizeof(BM)), 256);
    const size_t sched_b = (size_t)round_up((int64_t)((6 * B + 3) * 8), 256);
    const size_t stat_b = (size_t)round_up((int64_t)(B * sizeof(ItemStatus)), 256);
    const size_t one = desc_b + sched_b + stat_b;
    LINAC_TRY(stage(slot, one * nparts));
    char* h = (char*)h_buf[slot];
    char* dv = (char*)d_buf[slot];
    for (int i = 0; i < nparts; ++i) fill(act, parts[i], h + i * one, dv + i * one, desc_b, sched_b, out + i);
    LINAC_HIP(hipMemcpyAsync(dv, h, one * (nparts - 1) + desc_b + sched_b, hipMemcpyHostToDevice, s));
    return LINAC_OK;
}

template <class LT>
void Engine<LT>::fill(const std::vector<It*>& act, int part, char* h, char* dv, size_t desc_b, size_t sched_b,
                      Staged* out) {
    const size_t B = act.size();
    BM* hd = (BM*)h;
    int64_t* hs = (int64_t*)(h + desc_b);
    int64_t* ds = (int64_t*)(dv + desc_b);
    out->macs = out->bytes = 0;
    out->cols = 1;
    for (size_t m = 0; m < B; ++m) {
        BM d = act[m]->d;
        if (part == 4) {  // compact view of the candidate block (N = 0: not on the fast path)
            BM c = act[m]->dc;
            c.ld = XLD;
            c.N = sel[m] ? d.bcur : 0;
            c.M = d.bcur;
            c.f = d.f;
            c.pd = d.pd;
            c.pinv = d.pinv;
            c.J = d.J;
            c.bcur = d.bcur;
            c.e0 = d.e0;
            c.r0 = 0;
            d = c;
        } else if (part == 5) {  // Z = -P B into the panel buffer's Z columns (nT = 0: not fast)
            d.A = d.X + BO;
            d.T0 = 0;
            d.ld = XLD;
            d.nT = sel[m] ? d.bcur : 0;
            d.kA = 0;
            d.KqA = 0;
            d.Kq = (int32_t)round_up(d.bcur, 64);
            d.planesL = d.planesL1;
            d.strideL = d.strideL1;
            d.planesB = d.planesB1;
            d.strideB = STRIDE_W;
        } else if (part != 0) {
            // the next panel's n1 columns, then (composite updates only) the rest
            // in two halves; the first half covers at least the panel after next
            const int64_t n1 = std::min<int64_t>(BO, d.nT), rest = d.nT - n1;
            const int64_t h = rest <= BO ? rest : std::min<int64_t>(rest, std::max<int64_t>(BO, round_up(rest / 2, 64)));
            const bool comp = d.kA > 0;
            if (!comp) {  // a deferring, plain or last panel: its own left operand
                d.planesL = d.planesL1;
                d.strideL = d.strideL1;
            }
            if (part == 1) {
                d.nT = n1;
                if (!comp) {
                    d.planesB = d.planesB1;
                    d.strideB = STRIDE_W;
                }
            } else {
                const int64_t off = part == 2 ? n1 : n1 + h;
                d.T0 += off;
                d.nT = !(comp || d.full) ? 0 : (part == 2 ? h : rest - h);
                d.planesB += off * d.Kq;
            }
            const int64_t K = d.kA + d.q0;
            if (K > 0 && d.nT > 0) {
                // modular MACs of the rows that need the update; bytes: C read
                // and written plus both operands' limb planes once
                out->macs += (double)std::max<int64_t>(0, d.N - d.r0 - K) * K * d.nT;
                out->bytes += 2.0 * sizeof(typename LT::T) * (d.N - d.r0) * d.nT +
                              (double)LT::NL * d.Kq * ((d.N - d.r0) + d.nT);
                out->cols = std::max(out->cols, d.nT);
            }
        }
        hd[m] = d;
    }
    // GEMM tile width: 64-bit trailing parts take the tuned width (knob
    // tile64), the fast path's Z GEMM 32
    int64_t nmax = 1;  // the step's size class: its largest matrix (not the view of this part)
    for (size_t m = 0; m < B; ++m) nmax = std::max(nmax, std::max(act[m]->d.N, act[m]->d.M));
    const int ni64 = (part >= 1 && part <= 3 && tuning().pick(Tuning::TILE64, nmax) == 2) ? 2 : 1;
    const int64_t TN = std::is_same<LT, LimbsU64>::value ? 32 * ni64 : G_N;
    out->ni64 = ni64;
    // 32-bit inner GEMM: one workgroup per 64-row stripe, or (few stripes)
    // up to four that share a stripe's column tiles
    int64_t stripes = 0;
    for (size_t m = 0; m < B; ++m) stripes += ceil_div(hd[m].N, G_M);
    // (64-bit tiles are half as wide and their fold is heavier: up to eight per stripe)
    // (knob inner_split: the cap, 0 = 8 for 64-bit and 4 for 32-bit storage)
    // (a step of 8 or more matrices already gives the inner GEMM workgroups
    // enough: at most 4 per stripe, which also spares the left-operand capture
    // of split stripes; measured 3% faster on 64 matrices of 3000^2)
    const int64_t cap_class = tuning().pick(Tuning::INNER_SPLIT, nmax);
    const int64_t cap_knob = B >= 8 && cap_class > 4 ? 4 : cap_class;
    const int split = std::is_same<LT, LimbsU64>::value
                          ? (int)std::max<int64_t>(1, std::min<int64_t>(cap_knob ? cap_knob : 8,
                                                                        (cap_knob ? 64 * cap_knob : 512) /
                                                                            std::max<int64_t>(1, stripes)))
                          : (int)std::max<int64_t>(1, std::min<int64_t>(cap_knob ? cap_knob : 4,
                                                                        (cap_knob ? 64 * cap_knob : 256) /
                                                                            std::max<int64_t>(1, stripes)));
    out->inner_split = split;
    int64_t* pa = hs;
    int64_t* ma = pa + (B + 1);
    int64_t* pb = ma + B;
    int64_t* mb = pb + (B + 1);
    int64_t* pc = mb + B;
    int64_t* mc = pc + (B + 1);
    pa[0] = pb[0] = pc[0] = 0;
    for (size_t m = 0; m < B; ++m) {
        const BM& d = hd[m];
        ma[m] = ceil_div(d.N, G_M);
        // inner GEMM: split workgroups per 64-row stripe
        pa[m + 1] = pa[m] + ((part == 0 || part == 4) ? ma[m] * split : 0);
        // (the trailing and composite GEMMs cover rows [r0, N))
        const int64_t rt = ceil_div(d.N - d.r0, G_M);
        const bool live = part != 0 && part != 4 && d.Kq > 0 && d.nT > 0 && rt > 0;
        const bool wlive = part == 1 && d.kA > 0 && d.q0 > 0 && d.nT > 0 && rt > 0;
        mb[m] = std::max<int64_t>(1, rt);
        pb[m + 1] = pb[m] + (wlive ? rt * ceil_div(d.kA, TN) : 0);
        mc[m] = std::max<int64_t>(1, rt);
        pc[m + 1] = pc[m] + (live ? rt * ceil_div(d.nT, TN) : 0);
    }
    // direct-to-LDS staging: 16-byte aligned C rows, operand offsets below 2^31
    bool offs = true, aligned = true;
    for (size_t m = 0; m < B; ++m) {
        const BM& d = hd[m];
        offs = offs && prows(d.N) * (int64_t)d.Kq < (1ll << 31) && prows(d.nT) * (int64_t)d.Kq < (1ll << 31);
        aligned = aligned && ((((uintptr_t)(d.A + d.T0)) | (uintptr_t)(d.ld * (int64_t)sizeof(typename LT::T))) & 15) == 0;
    }
    out->lds_main = offs && aligned;
    out->lds_wq = offs;
    out->desc = (BM*)dv;
    const int grp = (int)tuning().pick(Tuning::GEMM_GROUP, nmax);
    out->inner = TileSched{ds, ds + (B + 1), (int)B, grp, 0, -1};
    out->wq = TileSched{ds + (B + 1) + B, ds + 2 * (B + 1) + B, (int)B, grp, 0, -1};
    out->main = TileSched{ds + 2 * (B + 1) + 2 * B, ds + 3 * (B + 1) + 2 * B, (int)B, grp, 0, -1};
    out->inner_tiles = pa[B];
    out->wq_tiles = pb[B];
    out->main_tiles = pc[B];
    out->status = (ItemStatus*)(dv + desc_b + sched_b);
    out->status_h = (ItemStatus*)(h + desc_b + sched_b);  // pinned; never part of an upload
}

static unsigned gs_grid(int64_t work, int64_t cap = 4096) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cap, ceil_div(work, 256)));
}

template <class LT>
int Engine<LT>::init_items(const std::vector<It*>& items, hipStream_t s) {
    if (items.empty()) return LINAC_OK;
    Staged st;
    LINAC_TRY(upload(items, s, 0, 0, &st));
    int64_t maxw = 1;
    for (auto* it : items) maxw = std::max<int64_t>(maxw, it->d.N * it->d.M);
    k_item_init<LT><<<dim3(gs_grid(maxw), (unsigned)items.size()), 256, 0, s>>>(st.desc);
    LINAC_HIP(hipGetLastError());
    LINAC_HIP(hipStreamSynchronize(s));  // the next upload reuses this staging slot
    return LINAC_OK;
}

// The full-rank panel fast path for the items whose panel can be full rank
// (at least bcur unused rows): inner steps on the compact candidate blocks,
// one status read, then for every block of full rank the GEMM Z = -P B and
// the commit.  `fast[m]` tells which items are done with this panel's inner
// phase.  Runs on s after the panel load.
template <class LT>
int Engine<LT>::fast_panel(const std::vector<It*>& act, hipStream_t s, const Staged& in, const Staged& cs,
                           const Staged& ws, int64_t ninner, int64_t maxN, std::vector<char>& fast, Profiler* prof) {
    constexpr bool wide = std::is_same<LT, LimbsU64>::value;
    const unsigned B = (unsigned)act.size();
    fast.assign(B, 0);
    // (the candidate rows were gathered by the panel load's last workgroup)
    k_cand_copy<LT><<<dim3(BO / 2, B), 256, 0, s>>>(in.desc, cs.desc);
    const int mu_passes = u_passes(B);
    const unsigned mu_grid = (unsigned)(MU_BLOCKS / mu_passes +
                                        ((wide || cs.inner_split > 1) ? std::min<int64_t>(4096, ceil_div(BO, MU_ROWS)) : 0));
    const unsigned igrid = (unsigned)(ceil_div(cs.inner_tiles, NUM_XCD) * NUM_XCD);
    if constexpr (!wide) {
        if (tuning().fast_panel == 1) {  // the block's inner steps in one workgroup per matrix
            k_block_factor<<<B, 64 * FB_WAVES, 0, s>>>(cs.desc, epoch);
            epoch += (int)ninner;
        }
    }
    if (wide || tuning().fast_panel == 2)  // the inner-step kernels on the block
        for (int64_t i = 0; i < ninner; ++i) {
            launch_inner_factor<LT>(B, s, cs.desc, epoch);
            k_make_u<LT><<<dim3(mu_grid, B), dim3(64, 4), 0, s>>>(cs.desc, mu_passes);
            if constexpr (wide) k_gemm_inner64<<<igrid, G_T, 0, s>>>(cs.desc, cs.inner, epoch, cs.inner_split);
            else k_gemm_inner<<<igrid, G_T, 0, s>>>(cs.desc, cs.inner, epoch, cs.inner_split);
            ++epoch;
        }
    // The Z GEMM and the commit are queued for every selected matrix before
    // the host reads whether its block reached full rank (k_fast_gate drops
    // the others on the device): the status read-back overlaps these kernels
    // instead of holding the GPU idle for a host round trip.
    k_fast_gate<LT><<<1, 256, 0, s>>>(ws.desc, cs.desc, const_cast<int64_t*>(ws.main.pre), (int)B, cs.status,
                                      in.desc);
    int64_t maxR = 1;  // rows from the row offset on (the part-5 descriptors carry the items' current r0)
    for (auto* it : act) maxR = std::max<int64_t>(maxR, it->d.N - it->d.r0);
    const int nl = (int)gs_grid(maxR * (BO / 4), 16384);
    const int nb = (int)ceil_div((int64_t)BO * BO, 256);
    k_fast_operands<LT><<<dim3(nl + nb, B), 256, 0, s>>>(ws.desc, cs.desc, nl);
    // (the status copy after the operands: off the chain to the Z GEMM; the
    // host still has it well before the tail ends)
    LINAC_HIP(hipMemcpyAsync(cs.status_h, cs.status, B * sizeof(ItemStatus), hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipEventRecord(ev_status, s));
    const unsigned grid = (unsigned)(ceil_div(ws.main_tiles, NUM_XCD) * NUM_XCD);
    // (profiled as a launch of the trailing-update kernel; its work is known
    // once the status says which blocks were of full rank)
    const size_t zsample = prof ? prof->samples.size() : 0;
    if (prof) LINAC_TRY(prof->sample_begin(s, act[0]->d.J, 0.0, 0.0));
    if constexpr (wide) k_gemm_outer64<0><<<grid, G_T, 0, s>>>(ws.desc, ws.main);
    else launch_outer32<0>(grid, s, ws.desc, ws.main, ws.lds_main);
    if (prof) LINAC_TRY(prof->sample_end(s));
    k_fast_commit<LT><<<dim3(BO / 2 + 1, B), 256, 0, s>>>(in.desc, cs.desc, ws.desc, act[0]->d.e0);
    LINAC_HIP(hipGetLastError());
    const auto spin0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t q = hipEventQuery(ev_status);
        if (q == hipSuccess) break;
        if (q != hipErrorNotReady) return fail(LINAC_ERR_DEVICE, "hipEventQuery: %s", hipGetErrorString(q));
        __builtin_ia32_pause();
    }
    spin_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - spin0).count();
    lo_fast.assign(B, 0);
    for (unsigned m = 0; m < B; ++m) {
        const BM& d = act[m]->d;
        fast[m] = sel[m] && cs.status_h[m].q0 == d.bcur && cs.status_h[m].rank == d.bcur ? 1 : 0;
        lo_fast[m] = cs.status_h[m].lo;
        if (prof && fast[m]) {
            // Z = -P B for the rows from r0 on outside the block (the block's
            // rows get theirs from it): MACs as for a trailing update of depth
            // and width bcur; bytes: Z read and written, both planes once
            const double rows = (double)(d.N - d.r0), kq = (double)round_up(d.bcur, 64);
            prof->samples[zsample].macs += std::max(0.0, rows - d.bcur) * d.bcur * d.bcur;
            prof->samples[zsample].bytes += 2.0 * sizeof(typename LT::T) * rows * d.bcur + LT::NL * kq * (rows + d.bcur);
        }
    }
    return LINAC_OK;
}

template <class LT>
int Engine<LT>::step(const std::vector<It*>& items, hipStream_t s, Profiler* prof) {
    constexpr bool wide = std::is_same<LT, LimbsU64>::value;
    const auto step0 = std::chrono::steady_clock::now();
    spin_us = 0;
    std::vector<It*> act;
    for (auto* it : items)
        if (!it->finished) act.push_back(it);
    if (act.empty()) return LINAC_OK;
    LINAC_TRY(ensure_stream());
    const unsigned B = (unsigned)act.size();
    int64_t maxN = 1, ninner = 1;
    for (auto* it : act) {
        BM& d = it->d;
        d.bcur = (int32_t)std::min<int64_t>(BO, d.M - d.J);
        d.rank0 = it->rank;
        d.e0 = epoch;
        d.q0 = 0;
        d.Kq = 0;
        d.nT = 0;
        maxN = std::max(maxN, d.N);
        ninner = std::max<int64_t>(ninner, ceil_div(d.bcur, BI));
    }
    // Trailing parts due now run in slices, one beside each inner step's
    // factorisation and U; the step's inner GEMM, which needs the whole GPU,
    // waits for its slice.
    const int64_t tag = act[0]->d.J;
    int64_t due_tiles = 0;
    for (auto& b : bulk)
        if (b.due <= step_no) due_tiles += b.t.main_tiles - b.next;
    const bool slicing = due_tiles > 0 && tuning().lookahead != 0 && tuning().timeslice != 0;
    const int64_t per_slice = slicing ? ceil_div(due_tiles, std::min<int64_t>(ninner, NEV)) : 0;
    // which items may take the full-rank fast path (at least bcur unused rows)
    sel.assign(B, 0);
    bool any_fast = false;
    if (tuning().fast_panel != 0 && !slicing)
        for (unsigned m = 0; m < B; ++m) {
            const BM& d = act[m]->d;
            sel[m] = d.N - act[m]->rank >= d.bcur ? 1 : 0;
            any_fast = any_fast || sel[m];
        }
    // --- inner phase on s, after the previous step's first trailing part; its
    // descriptors, with the fast path's compact views (part 4) and Z-GEMM
    // descriptors (part 5), in one copy (which need not wait for the event)
    Staged st3[3];
    const int parts[3] = {0, 4, 5};
    LINAC_TRY(upload_list(act, s, 0, parts, any_fast ? 3 : 1, st3));
    const Staged& in = st3[0];
    if (ev_pending) LINAC_HIP(hipStreamWaitEvent(s, ev_ready, 0));
    int64_t maxRX = 1;  // panel-buffer entries from the row offset on
    for (auto* it : act) maxRX = std::max<int64_t>(maxRX, (it->d.N - it->d.r0) * XLD);
    const unsigned lgrid = gs_grid(maxRX / 4, 8192);
    if (any_fast)  // the last workgroup gathers the candidate rows
        k_panel_load<LT><<<dim3(lgrid + 1, B), 256, 0, s>>>(in.desc, first_step ? 1 : 0, st3[1].desc);
    else
        k_panel_load<LT><<<dim3(lgrid, B), 256, 0, s>>>(in.desc, first_step ? 1 : 0);
    first_step = false;
    // (the 32-bit inner GEMM captures its left operand itself)
    const bool planes_a = wide || in.inner_split > 1;
    const int mu_passes = u_passes(B);
    const unsigned mu_grid =
        (unsigned)(MU_BLOCKS / mu_passes + (planes_a ? std::min<int64_t>(4096, ceil_div(maxN, MU_ROWS)) : 0));
    auto group = [&](int64_t i) -> int {
        const bool sl = slicing && i < NEV;
        if (sl) {
            LINAC_HIP(hipEventRecord(ev_f[i], s));
            LINAC_HIP(hipStreamWaitEvent(so, ev_f[i], 0));
            int64_t want = per_slice;
            for (auto& b : bulk) {
                if (want <= 0 || b.due > step_no) break;
                const int64_t before = b.next;
                LINAC_TRY(launch_slice(b, want, prof, tag));
                want -= b.next - before;
            }
            LINAC_HIP(hipEventRecord(ev_c[i], so));
        }
        launch_inner_factor<LT>(B, s, in.desc, epoch);
        k_make_u<LT><<<dim3(mu_grid, B), dim3(64, 4), 0, s>>>(in.desc, mu_passes);
        if (sl) LINAC_HIP(hipStreamWaitEvent(s, ev_c[i], 0));
        const unsigned grid = (unsigned)(ceil_div(in.inner_tiles, NUM_XCD) * NUM_XCD);
        if constexpr (wide) k_gemm_inner64<<<grid, G_T, 0, s>>>(in.desc, in.inner, epoch, in.inner_split);
        else k_gemm_inner<<<grid, G_T, 0, s>>>(in.desc, in.inner, epoch, in.inner_split);
        ++epoch;
        LINAC_HIP(hipGetLastError());
        return LINAC_OK;
    };
    // full-rank panels through their candidate blocks; the regular inner steps
    // over every row for the rest
    std::vector<char> fast(B, 0);
    int nfast = 0;
    if (any_fast) {
        LINAC_TRY(fast_panel(act, s, in, st3[1], st3[2], ninner, maxN, fast, prof));
        for (char f : fast) nfast += f;
    }
    std::vector<ItemStatus> st(B);
    if (nfast < (int)B)
        for (int64_t i = 0; i < ninner; ++i) LINAC_TRY(group(i));
    // one synchronisation per step; extra groups only for rank-deficient panels
    for (int64_t guard = 0; nfast < (int)B; ++guard) {
        // into pinned memory, then a spin on the event: the next launches are
        // on the critical path whenever the inner phase outlasts the trailing
        // update, and a blocking synchronisation wakes up late
        k_gather_status<LT><<<B, 1, 0, s>>>(in.desc, in.status);
        LINAC_HIP(hipMemcpyAsync(in.status_h, in.status, B * sizeof(ItemStatus), hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipEventRecord(ev_status, s));
        for (;;) {
            const hipError_t q = hipEventQuery(ev_status);
            if (q == hipSuccess) break;
            if (q != hipErrorNotReady)
                return fail(LINAC_ERR_DEVICE, "hipEventQuery: %s", hipGetErrorString(q));
            __builtin_ia32_pause();
        }
        std::copy(in.status_h, in.status_h + B, st.begin());
        int64_t behind = 0;
        for (unsigned m = 0; m < B; ++m)
            if (st[m].c0_next < st[m].bcur) behind = std::max<int64_t>(behind, ceil_div(st[m].bcur - st[m].c0_next, BI));
        if (behind == 0) break;
        if (guard > (int64_t)BO * (BI + 1)) return fail(LINAC_ERR_DEVICE, "blocked elimination did not converge");
        for (int64_t i = 0; i < behind; ++i) LINAC_TRY(group(NEV));
    }
    if (nfast == (int)B) {
        // every panel was full rank: the status is known without reading it,
        // and the trailing update waits for the inner phase by an event
        for (unsigned m = 0; m < B; ++m) {
            const BM& d = act[m]->d;
            st[m] = ItemStatus{d.bcur, d.bcur, d.bcur, 0, act[m]->rank + d.bcur, lo_fast[m]};
        }
        LINAC_HIP(hipEventRecord(ev_inner, s));
    }
    for (unsigned m = 0; m < B; ++m) act[m]->rank = st[m].rank;
    // whatever of the parts due by now is left (rank-deficient panels add
    // groups; slices round), before this step's own trailing update and
    // before the copy below refills the staging slot of the parts two steps
    // back (never the case on the fast path, which runs without slices: its
    // parts due now were launched at the end of the previous step)
    LINAC_TRY(flush_due(step_no, prof, tag));
    // --- trailing update on so: first the next panel's columns (then the
    // panel goes back to A and the next inner phase may start), then the rest.
    // The columns beyond the next panel are either updated together with the
    // previous panel's deferred part (kA > 0: one composite GEMM) or deferred
    // to the next step (save_w).
    for (unsigned m = 0; m < B; ++m) {
        It* it = act[m];
        BM& d = it->d;
        d.q0 = st[m].q0;
        d.rank_start = st[m].rank - st[m].q0;
        // rows above the first unused row at this panel's load are pivot rows
        // of earlier panels; with forward elimination they leave the updates
        // (in whole 64-row tiles).  A composite step keeps the deferred
        // panel's r0: both panels' updates then cover the same rows.
        if (d.kA == 0) it->r0 = it->forward ? std::min<int64_t>(st[m].lo, d.N) / G_M * G_M : 0;
        d.r0 = it->r0;
        if (d.r0 > 0 && d.q0 > 0) it->echelon.push_back(EchelonPanel{d.J, d.bcur, d.r0});
        if (debug_steps())
            fprintf(stderr, "step J=%lld lo=%lld r0=%lld kA=%d q0=%d rank=%lld fast=%d host %.1f us spin %.1f us\n",
                    (long long)d.J, (long long)st[m].lo, (long long)d.r0, d.kA, d.q0, (long long)st[m].rank,
                    (int)fast[m],
                    std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - step0).count(),
                    spin_us);
        d.T0 = d.J + d.bcur;
        d.nT = d.M - d.T0;
        d.KqA = (int32_t)round_up(d.kA, 64);
        d.Kq = d.KqA + (int32_t)round_up(d.q0, 64);
        d.full = st[m].rank >= d.N ? 1 : 0;  // no row left to pivot: the remaining columns are final after this step
        d.store_now = (d.J + d.bcur >= d.M || d.full) ? 1 : 0;  // the item's last step
        d.bst = 0;  // (the panel load of this step has written the previous panel back)
        d.save_w = (!d.full && d.kA == 0 && d.q0 > 0 && d.nT > BO) ? 1 : 0;
    }
    const int p1 = 1 + parity;
    parity ^= 1;
    Staged tp[3];
    // (the descriptor copy goes ahead of the wait for the fast path's tail
    // kernels, which do not read it)
    LINAC_TRY(upload(act, so, p1, 1, tp, 3));
    const auto t_up = std::chrono::steady_clock::now();
    if (nfast == (int)B) LINAC_HIP(hipStreamWaitEvent(so, ev_inner, 0));
    const Staged &t1 = tp[0], &t2 = tp[1], &t3 = tp[2];
    bool operands = t1.main_tiles > 0;
    for (auto* it : act) operands = operands || it->d.save_w;
    if (operands) {
        // -W of the current panel (kept in W when deferred) beside, for
        // composite items, the planes of W[Q']; then W <- W + W' W[Q']; then -W
        // of the composite beside the right operand of the next panel's update
        int64_t maxR = 1;  // rows from this step's row offset on
        for (auto* it : act) maxR = std::max<int64_t>(maxR, it->d.N - it->d.r0);
        const int nl = (int)gs_grid(maxR * (BO / 4), 16384);
        const bool comp = t1.wq_tiles > 0, main1 = t1.main_tiles > 0;
        constexpr int RW = BO / 64;  // right-operand tiles of W[Q'] per dimension
        const int rx = (int)std::min<int64_t>(2048, ceil_div(t1.cols, 64));
        const int nr = main1 ? rx * (2 * BO / 64) : 0;
        // (a deferred panel followed by one without pivots still needs its
        // composite's planes, though there is no W' W[Q'] to add)
        bool deferred = false;
        for (auto* it : act) deferred = deferred || it->d.kA > 0;
        if (!comp && !deferred) {
            // no composite: -W beside the right operand, one launch
            k_make_ops<LT><<<dim3(nl + nr, B), 256, 0, so>>>(t1.desc, 0, epoch, nl, main1 ? 0 : -1, rx);
        } else {
            k_make_ops<LT><<<dim3(nl + (comp ? RW * RW : 0), B), 256, 0, so>>>(t1.desc, 0, epoch, nl, comp ? 1 : -1,
                                                                              RW);
            if (comp) {
                const unsigned grid = (unsigned)(ceil_div(t1.wq_tiles, NUM_XCD) * NUM_XCD);
                if constexpr (wide) launch_outer64<1>(grid, so, t1.desc, t1.wq, t1.ni64);
                else launch_outer32<1>(grid, so, t1.desc, t1.wq, t1.lds_wq);
            }
            k_make_ops<LT><<<dim3(nl + nr, B), 256, 0, so>>>(t1.desc, 1, epoch, nl, main1 ? 0 : -1, rx);
        }
        LINAC_HIP(hipGetLastError());
    }
    LINAC_TRY(trailing(t1, prof, tag, operands));
    // The panel goes back to A on the inner stream, beside the rest of the
    // update: it writes only the panel's own columns, which no later step
    // changes, and reads the panel buffer, which only the next inner phase
    // (after it on that stream) overwrites.
    LINAC_HIP(hipEventRecord(ev_ready, so));
    // An item's last panel is written back here; the others' by the next
    // panel load, from the same buffer rows it then refills.  (The next step
    // waits for ev_ready after its descriptor copy, which needs no wait.)
    int64_t maxpanel = 0;  // (one row per workgroup pass, rows from the row offset on)
    for (auto* it : act)
        if (it->d.store_now) maxpanel = std::max<int64_t>(maxpanel, (it->d.N - it->d.r0) * it->d.bcur);
    if (maxpanel > 0) {
        LINAC_HIP(hipStreamWaitEvent(s, ev_ready, 0));
        k_panel_store<LT><<<dim3(gs_grid(maxpanel, 8192), B), 256, 0, s>>>(t1.desc);
    }
    LINAC_HIP(hipGetLastError());
    // The next inner phase waits only for the above.  The composite update's
    // first half is due beside the next step's inner phase, its second half
    // beside the one after (its columns lie beyond the next panel's, which
    // that step's own update needs first).  Without time slices each part
    // runs whole, queued behind the next-panel update of the step before
    // its due step; without look-ahead everything runs now.
    const bool lookahead = tuning().lookahead != 0;
    if (t2.main_tiles > 0) bulk.push_back(Bulk{t2, 0, step_no + 1});
    if (t3.main_tiles > 0) bulk.push_back(Bulk{t3, 0, step_no + 2});
    if (!lookahead) LINAC_TRY(flush_due(INT64_MAX, prof, tag));
    else if (tuning().timeslice == 0) LINAC_TRY(flush_due(step_no + 1, prof, tag));
    if (!lookahead) LINAC_HIP(hipEventRecord(ev_ready, so));  // the next inner phase waits for all of it
    if (debug_steps())
        fprintf(stderr, "  upload at %.1f us, step end at %.1f us\n",
                std::chrono::duration<double, std::micro>(t_up - step0).count(),
                std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - step0).count());
    ++step_no;
    ev_pending = true;  // the next panel load waits for ev_ready (the next panel's columns, or all parts)
    for (auto* it : act) {
        it->d.kA = it->d.save_w ? it->d.q0 : 0;  // deferred to the next step
        if (!it->d.store_now) {  // the next panel load writes this panel back
            it->d.Jst = it->d.J;
            it->d.bst = it->d.bcur;
        }
        it->d.J += it->d.bcur;
        if (it->d.J >= it->d.M || it->d.full) it->finished = true;
    }
    return LINAC_OK;
}

template <class LT>
static void describe(BMat<LT>& d, typename LT::T* A, int64_t ld, int64_t N, int64_t M, const typename LT::F& f) {
    d.A = A;
    d.ld = ld;
    d.N = N;
    d.M = M;
    d.f = f;
    d.pd = (double)f.p;
    d.pinv = 1.0 / (double)f.p;
    d.J = 0;
    d.kA = 0;
}

template <class LT>
static void print_timers(const BMat<LT>& d, DeviceContext* ctx) {
    unsigned long long tp[8];
    if (hipMemcpy(tp, d.bs->tphase, sizeof(tp), hipMemcpyDeviceToHost) != hipSuccess || tp[5] == 0) return;
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess) return;
    const double us = 1e3 / (double)rate_khz;
    fprintf(stderr, "k_inner_factor phases over %llu calls (us/call): scan %.2f load %.2f gj %.2f norm %.2f out %.2f\n",
            tp[5], tp[0] * us / tp[5], tp[1] * us / tp[5], tp[2] * us / tp[5], tp[3] * us / tp[5],
            tp[4] * us / tp[5]);
}

// The context's engine for storage width LT (created on first use; freed by
// the shutdown hook with the context).
template <class LT>
static void free_engine(void* e) {
    delete static_cast<Engine<LT>*>(e);
}

static void free_cached_engines(DeviceContext* ctx) {
    if (ctx->engines[0]) free_engine<LimbsU32>(ctx->engines[0]);
    if (ctx->engines[1]) free_engine<LimbsU64>(ctx->engines[1]);
    ctx->engines[0] = ctx->engines[1] = nullptr;
}

template <class LT>
static Engine<LT>* cached_engine(DeviceContext* ctx) {
    void*& slot = ctx->engines[std::is_same<LT, LimbsU64>::value ? 1 : 0];
    if (!slot) {
        add_shutdown_hook(free_cached_engines);
        slot = new Engine<LT>();
    }
    return static_cast<Engine<LT>*>(slot);
}

// Single matrix through the engine (a batch of one); workspace from the
// context's cached scratch, elimination state from `ws`.
template <class LT>
static int blocked_run(typename LT::T* A, int64_t ld, int64_t N, int64_t M, const typename LT::F& f, hipStream_t s,
                       const ElimWorkspace& ws, DeviceContext* ctx, Profiler* prof,
                       std::vector<EchelonPanel>* echelon) {
    const WorkspaceLayout w = workspace_layout<LT>(N, M, false);
    void* p = nullptr;
    LINAC_TRY(ctx->get(SLOT_PANEL_A, w.total, &p));
    Item<LT> it;
    describe(it.d, A, ld, N, M, f);
    carve(it.d, (char*)p, w, N, M, false);
    carve_compact(it.dc, (char*)p, w);
    it.d.st = ws.st;
    it.d.piv_rows = ws.piv_rows;
    it.d.piv_cols = ws.piv_cols;
    it.d.used = ws.used;
    it.forward = echelon != nullptr && tuning().echelon != 0;
    Engine<LT>& eng = *cached_engine<LT>(ctx);
    eng.reset_for_solve();
    std::vector<Item<LT>*> items{&it};
    // state and row flags were initialised by the caller; epochs here
    k_epoch_init<<<gs_grid(N), 256, 0, s>>>(it.d.epoch_of, N);
    const bool timing = getenv("LINAC_DEBUG_TIMING") != nullptr;
    if (timing) {  // the phase counters live in reused scratch
        LINAC_HIP(hipMemsetAsync(it.d.bs->tphase, 0, sizeof(it.d.bs->tphase), s));
        LINAC_HIP(hipMemsetAsync(it.dc.bs->tphase, 0, sizeof(it.dc.bs->tphase), s));
    }
    int rc = LINAC_OK;
    while (!it.finished && rc == LINAC_OK) rc = eng.step(items, s, prof);
    const int rd = eng.drain();
    LINAC_TRY(rc);
    LINAC_TRY(rd);
    if (echelon) *echelon = std::move(it.echelon);
    LINAC_HIP(hipStreamSynchronize(s));
    if (timing) {
        print_timers(it.d, ctx);
        print_timers(it.dc, ctx);  // inner steps on the candidate blocks of full-rank panels
        unsigned long long tp[8];
        int rate_khz = 0;
        if (hipMemcpy(tp, it.dc.bs->tphase, sizeof(tp), hipMemcpyDeviceToHost) == hipSuccess && tp[6] &&
            hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess) {
            const double us = 1e3 / (double)rate_khz;
            fprintf(stderr, "k_block_factor phases over %llu blocks (us/block): gj %.1f norm %.1f xr %.1f u %.1f uw %.1f upd %.1f\n",
                    tp[6], tp[0] * us / tp[6], tp[1] * us / tp[6], tp[2] * us / tp[6], tp[3] * us / tp[6],
                    tp[4] * us / tp[6], tp[5] * us / tp[6]);
        }
        (void)hipMemset(it.dc.bs->tphase, 0, sizeof(tp));
    }
    return LINAC_OK;
}

template <class F>
int blocked_dispatch(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                     const ElimWorkspace& ws, DeviceContext* ctx, bool forced, Profiler* prof,
                     std::vector<EchelonPanel>* echelon) {
    if constexpr (std::is_same<F, FieldU32>::value) {
        if (f.p > LIMB_P_MAX) return NOT_APPLICABLE;
        if (!forced && (N < tuning().gf_blocked_min_rows || M < tuning().gf_blocked_min_cols)) return NOT_APPLICABLE;
        return blocked_run<LimbsU32>(A, ld, N, M, f, s, ws, ctx, prof, echelon);
    } else if constexpr (std::is_same<F, FieldU64>::value) {
        if (f.p < LIMB64_P_MIN || f.p >= LIMB64_P_MAX) return NOT_APPLICABLE;
        if (!forced && (N < tuning().gf_blocked_min_rows || M < tuning().gf_blocked_min_cols)) return NOT_APPLICABLE;
        return blocked_run<LimbsU64>(A, ld, N, M, f, s, ws, ctx, prof, echelon);
    } else {
        (void)A; (void)ld; (void)N; (void)M; (void)f; (void)s; (void)ws; (void)ctx; (void)forced; (void)prof;
        (void)echelon;
        return NOT_APPLICABLE;
    }
}

template int blocked_dispatch<FieldU32>(uint32_t*, int64_t, int64_t, int64_t, const FieldU32&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*,
                                        std::vector<EchelonPanel>*);
template int blocked_dispatch<FieldU64>(uint64_t*, int64_t, int64_t, int64_t, const FieldU64&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*,
                                        std::vector<EchelonPanel>*);
template int blocked_dispatch<FieldGF2>(uint32_t*, int64_t, int64_t, int64_t, const FieldGF2&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, bool, Profiler*,
                                        std::vector<EchelonPanel>*);

// ---------------------------------------------------------------------------
// Exact modular GEMM with index lists (linac_gemm_sub_u32):
//   C[i, j] <- C[i, j] - sum_t A[i, a_cols[t]] B[b_rows[t], j]   (mod p)
// on the trailing-update kernel, in depth chunks of at most 2 BO = 512 (the
// limb fold's bound).  Every chunk's limb planes are built before the first
// chunk's GEMM, so C may be the same matrix as A or B.
// ---------------------------------------------------------------------------

constexpr int GS_KC = 2 * BO;
using BMat32 = BMat<LimbsU32>;

// planes[l][i][t] = digit l of A[row(i), cols[t0 + t]] for t < kc (cols
// NULL: column t0 + t; rows NULL: row(i) = i, else rows[i]), zero up to Kpad;
// i < m.  (The GEMM kernel computes C - L B, so L holds A itself here, where
// the trailing update holds -W.)
template <class LT>
__global__ __launch_bounds__(256) void k_planes_cols(const typename LT::T* __restrict__ A, int64_t lda,
                                                     const int64_t* __restrict__ rows,
                                                     const int64_t* __restrict__ cols, int64_t t0, int kc, int Kpad,
                                                     int64_t m, typename LT::T p, int8_t* __restrict__ planes,
                                                     int64_t stride) {
    constexpr int NL = LT::NL;
    const int64_t total = m * Kpad;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = x / Kpad;
        const int t = (int)(x - i * Kpad);
        int8_t dg[NL] = {};
        if (t < kc) {
            const int64_t c = cols ? cols[t0 + t] : t0 + t;
            const int64_t r = rows ? rows[i] : i;
            balanced_digits(centred(A[r * lda + c], p), dg);
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) planes[l * stride + ptile(i, t, Kpad)] = dg[l];
    }
}

// planes[l][j][t] = digit l of B[rows[t0 + t], j] for t < kc (rows NULL: row
// t0 + t), zero up to Kpad; columns j < n.  64 x 64 tiles through LDS
// (blockIdx.x strides over column tiles, blockIdx.y = 64-deep tile); the 16
// digits of a limb are packed into four words with shifts.
template <class LT>
__global__ __launch_bounds__(256) void k_planes_rows(const typename LT::T* __restrict__ B, int64_t ldb,
                                                     const int64_t* __restrict__ rows, int64_t t0, int kc, int Kpad,
                                                     int64_t n, typename LT::T p, int8_t* __restrict__ planes,
                                                     int64_t stride) {
    using S = typename LT::S;
    constexpr int NL = LT::NL;
    __shared__ S tile[64][65];
    const int tb = blockIdx.y * 64;
    for (int64_t nb = (int64_t)blockIdx.x * 64; nb < n; nb += (int64_t)gridDim.x * 64) {
        for (int x = threadIdx.x; x < 64 * 64; x += 256) {
            const int tt = x >> 6, nn = x & 63;
            const int t = tb + tt;
            const int64_t j = nb + nn;
            S v = 0;
            if (t < kc && j < n) v = centred(B[(rows ? rows[t0 + t] : t0 + t) * ldb + j], p);
            tile[tt][nn] = v;
        }
        __syncthreads();
        const int nn = threadIdx.x >> 2, tq = (threadIdx.x & 3) * 16;
        const int64_t j = nb + nn;
        if (j < n) {
            uint32_t wd[NL][4];
#pragma unroll
            for (int l = 0; l < NL; ++l)
#pragma unroll
                for (int q = 0; q < 4; ++q) wd[l][q] = 0u;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                int8_t dg[NL];
                balanced_digits(tile[tq + i][nn], dg);
#pragma unroll
                for (int l = 0; l < NL; ++l) wd[l][i >> 2] |= (uint32_t)(uint8_t)dg[l] << (8 * (i & 3));
            }
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                const v4i w = {(int)wd[l][0], (int)wd[l][1], (int)wd[l][2], (int)wd[l][3]};
                *reinterpret_cast<v4i*>(planes + l * stride + ptile(j, tb + tq, Kpad)) = w;
            }
        }
        __syncthreads();
    }
}

static int gemm_sub_u32(uint32_t* C, int64_t ldc, const uint32_t* A, int64_t lda, int64_t a_ncols,
                        const int64_t* a_cols, const uint32_t* B, int64_t ldb, int64_t b_nrows, const int64_t* b_rows,
                        int64_t m, int64_t n, int64_t k, uint32_t p, uint32_t flags) {
    if (m < 0 || n < 0 || k < 0 || a_ncols < 0 || b_nrows < 0) return fail(LINAC_ERR_INVALID, "negative size");
    if (m == 0 || n == 0 || k == 0) return LINAC_OK;
    if (!(flags & LINAC_DEVICE_PTR)) return fail(LINAC_ERR_INVALID, "C, A and B must be device matrices (LINAC_DEVICE_PTR)");
    if (!C || !A || !B) return fail(LINAC_ERR_INVALID, "NULL matrix");
    if (ldc < n || ldb < n || lda < a_ncols) return fail(LINAC_ERR_INVALID, "a row stride is below its row length");
    if (!is_prime_u64(p) || p == 2 || p > LIMB_P_MAX)
        return fail(LINAC_ERR_INVALID, "p = %u is not an odd prime <= %llu", p, (unsigned long long)LIMB_P_MAX);
    for (int64_t t = 0; t < k; ++t) {
        const int64_t c = a_cols ? a_cols[t] : t, r = b_rows ? b_rows[t] : t;
        if (c < 0 || c >= a_ncols) return fail(LINAC_ERR_INVALID, "A column %lld outside [0, %lld)", (long long)c, (long long)a_ncols);
        if (r < 0 || r >= b_nrows) return fail(LINAC_ERR_INVALID, "B row %lld outside [0, %lld)", (long long)r, (long long)b_nrows);
    }
    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    const FieldU32 f = make_field_u32(p);
    const int64_t nch = ceil_div(k, GS_KC);
    // plane regions per chunk (L: m x Kpad, B: n x Kpad, four limbs each)
    std::vector<int64_t> offL(nch), offB(nch), sL(nch), sB(nch), kp(nch);
    int64_t totL = 0, totB = 0;
    for (int64_t c = 0; c < nch; ++c) {
        kp[c] = round_up(std::min<int64_t>(GS_KC, k - c * GS_KC), 64);
        sL[c] = round_up(prows(m) * kp[c], 256);
        sB[c] = round_up(prows(n) * kp[c], 256);
        offL[c] = totL;
        offB[c] = totB;
        totL += 4 * sL[c];
        totB += 4 * sB[c];
    }
    void* pp = nullptr;
    LINAC_TRY(ctx->get(SLOT_PANEL_B, (size_t)(totL + totB), &pp));
    int8_t* planesL = (int8_t*)pp;
    int8_t* planesB = planesL + totL;
    const int64_t ntiles = ceil_div(m, G_M) * ceil_div(n, G_N);
    const size_t desc_b = (size_t)round_up(nch * (int64_t)sizeof(BMat32), 256);
    const size_t idx_b = (size_t)round_up(k * 8, 256);
    void* pc = nullptr;
    LINAC_TRY(ctx->get(SLOT_PANEL_C, desc_b + 256 + 2 * idx_b, &pc));
    BMat32* d_desc = (BMat32*)pc;
    int64_t* d_sched = (int64_t*)((char*)pc + desc_b);  // pre[0..1], mts[0]
    int64_t* d_acols = (int64_t*)((char*)pc + desc_b + 256);
    int64_t* d_brows = d_acols + idx_b / 8;
    std::vector<BMat32> hd(nch);
    for (int64_t c = 0; c < nch; ++c) {
        BMat32 d{};
        d.A = C;
        d.ld = ldc;
        d.N = m;
        d.M = n;
        d.f = f;
        d.pd = (double)p;
        d.pinv = 1.0 / (double)p;
        d.planesL = planesL + offL[c];
        d.strideL = sL[c];
        d.planesB = planesB + offB[c];
        d.strideB = sB[c];
        d.T0 = 0;
        d.nT = n;
        d.Kq = (int32_t)kp[c];
        hd[c] = d;
    }
    const int64_t hs[3] = {0, ntiles, ceil_div(m, G_M)};
    LINAC_HIP(hipMemcpyAsync(d_desc, hd.data(), nch * sizeof(BMat32), hipMemcpyHostToDevice, s));
    LINAC_HIP(hipMemcpyAsync(d_sched, hs, sizeof(hs), hipMemcpyHostToDevice, s));
    if (a_cols) LINAC_HIP(hipMemcpyAsync(d_acols, a_cols, k * 8, hipMemcpyHostToDevice, s));
    if (b_rows) LINAC_HIP(hipMemcpyAsync(d_brows, b_rows, k * 8, hipMemcpyHostToDevice, s));
    for (int64_t c = 0; c < nch; ++c) {
        const int64_t t0 = c * GS_KC;
        const int kc = (int)std::min<int64_t>(GS_KC, k - t0);
        k_planes_cols<LimbsU32><<<gs_grid(m * kp[c], 16384), 256, 0, s>>>(
            A, lda, nullptr, a_cols ? d_acols : nullptr, t0, kc, (int)kp[c], m, p, planesL + offL[c], sL[c]);
        k_planes_rows<LimbsU32><<<dim3((unsigned)std::min<int64_t>(2048, ceil_div(n, 64)), (unsigned)(kp[c] / 64)), 256, 0, s>>>(
            B, ldb, b_rows ? d_brows : nullptr, t0, kc, (int)kp[c], n, p, planesB + offB[c], sB[c]);
    }
    const TileSched ts{d_sched, d_sched + 2, 1, (int)tuning().pick(Tuning::GEMM_GROUP, std::max(m, n)), 0, -1};
    bool lds_ok = (((uintptr_t)C | (uintptr_t)(ldc * 4)) & 15) == 0;
    for (int64_t c = 0; c < nch; ++c) lds_ok = lds_ok && prows(m) * kp[c] < (1ll << 31) && prows(n) * kp[c] < (1ll << 31);
    for (int64_t c = 0; c < nch; ++c)
        launch_outer32<0>((unsigned)(ceil_div(ntiles, NUM_XCD) * NUM_XCD), s, d_desc + c, ts, lds_ok);
    LINAC_HIP(hipGetLastError());
    LINAC_HIP(hipStreamSynchronize(s));
    return LINAC_OK;
}

// ---------------------------------------------------------------------------
// Echelon completion.  With forward elimination, a pivot row stops taking the
// updates of later panels once it lies above their r0, so it still holds
// entries in their pivot columns.  Sorted by column, the pivots give the
// result rows: row k is pivot k, and panel l owns rows [s_l, s_l + q_l).  Any
// vector v of the row space is v = sum_k v[c_k] R[k] (R: the RREF, c_k: the
// pivot columns), so a pivot row v of panel j satisfies
//     R[k] = v - sum_{l > j} v[P_l] R[rows of panel l],
// and the pivot columns of R are the identity.  Only the non-pivot columns F
// need arithmetic.  Processing the panels from the last, the rows of panel l
// are final when its correction is applied to every earlier row:
//     G[0:s_l, F] -= A[piv_rows[0:s_l], P_l] G[s_l:s_l+q_l, F]
// with G the pivot rows' entries in F (result order): one exact limb GEMM of
// s_l x q_l x |F| per panel on the trailing-update kernel.  Rows that took
// panel l's update hold zeros in P_l, so the correction leaves them alone.
// Columns of F left of the panel are zero in its rows and are skipped.
// ---------------------------------------------------------------------------

// G[k, j] = A[rows[k], fcols[j]] for k < rank, j < nF.
template <class T>
__global__ __launch_bounds__(256) void k_ech_gather(const T* __restrict__ A, int64_t ld,
                                                    const int64_t* __restrict__ rows,
                                                    const int32_t* __restrict__ fcols, int64_t rank, int64_t nF,
                                                    T* __restrict__ G, int64_t ldg) {
    for (int64_t k = blockIdx.y; k < rank; k += gridDim.y) {
        const T* src = A + rows[k] * ld;
        for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nF; j += (int64_t)gridDim.x * 256)
            G[k * ldg + j] = src[fcols[j]];
    }
}

// A rewritten as the RREF: row k < rank holds 1 in its pivot column, 0 in
// every other pivot column (cmap[c] = -1 - pivot index) and G[k, cmap[c]] in a
// non-pivot column c; rows from rank on are zero.  vec: rows are 16-byte
// aligned, so whole vectors within a row are stored as such.
template <class T>
__global__ __launch_bounds__(256) void k_ech_write(T* __restrict__ A, int64_t ld, int64_t N, int64_t M,
                                                   const T* __restrict__ G, int64_t ldg,
                                                   const int32_t* __restrict__ cmap, int64_t rank, int vec) {
    constexpr int V = 16 / (int)sizeof(T);
    const int64_t nv = (M + V - 1) / V;
    for (int64_t k = blockIdx.y; k < N; k += gridDim.y) {
        T* row = A + k * ld;
        const bool live = k < rank;
        auto value = [&](int64_t c) -> T {
            if (!live) return (T)0;
            const int32_t m = cmap[c];
            return m < 0 ? (T)((int64_t)(-1 - m) == k ? 1 : 0) : G[k * ldg + m];
        };
        for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < nv; v += (int64_t)gridDim.x * 256) {
            const int64_t c0 = v * V;
            if (vec && c0 + V <= M) {
                Vec16<T> x;
#pragma unroll
                for (int e = 0; e < V; ++e) x.v[e] = value(c0 + e);
                v4i out;
                __builtin_memcpy(&out, &x, 16);
                gst<v4i>(row + c0, out);
            } else {
                for (int e = 0; e < V && c0 + e < M; ++e) row[c0 + e] = value(c0 + e);
            }
        }
    }
}

template <class F>
int echelon_finalize(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                     const ElimWorkspace& ws, DeviceContext* ctx, const std::vector<int64_t>& rows,
                     const std::vector<int64_t>& cols, bool resorted, const std::vector<EchelonPanel>& panels) {
    using T = typename F::T;
    constexpr bool u32 = std::is_same<F, FieldU32>::value, u64 = std::is_same<F, FieldU64>::value;
    if constexpr (!u32 && !u64) {
        return gf_finalize_apply(A, ld, N, M, f, s, ws, ctx, false, rows, cols, resorted);
    } else {
        using LT = typename std::conditional<u32, LimbsU32, LimbsU64>::type;
        using BM = BMat<LT>;
        constexpr int NL = LT::NL;
        constexpr int64_t TN = u32 ? G_N : Tile64<1>::TN;
        if (panels.empty()) return gf_finalize_apply(A, ld, N, M, f, s, ws, ctx, false, rows, cols, resorted);
        const int64_t rank = (int64_t)rows.size();
        if (resorted && rank > 0) {
            LINAC_HIP(hipMemcpyAsync(ws.piv_rows, rows.data(), rank * 8, hipMemcpyHostToDevice, s));
            LINAC_HIP(hipMemcpyAsync(ws.piv_cols, cols.data(), rank * 8, hipMemcpyHostToDevice, s));
        }
        std::vector<int32_t> cmap((size_t)M, 0), fcols;
        for (int64_t k = 0; k < rank; ++k) cmap[cols[k]] = (int32_t)(-1 - k);
        std::vector<char> piv((size_t)M, 0);
        for (int64_t k = 0; k < rank; ++k) piv[cols[k]] = 1;
        fcols.reserve((size_t)(M - rank));
        for (int64_t c = 0; c < M; ++c)
            if (!piv[c]) {
                cmap[c] = (int32_t)fcols.size();
                fcols.push_back((int32_t)c);
            }
        const int64_t nF = (int64_t)fcols.size();
        const int64_t ldg = round_up(std::max<int64_t>(nF, 1), 64);
        // the GEMMs, last panel first
        struct Job {
            int64_t s, q, f0, Kpad;
        };
        std::vector<Job> jobs;
        int64_t maxL = 0, maxB = 0;
        if (nF > 0 && rank > 0)
            for (auto it = panels.rbegin(); it != panels.rend(); ++it) {
                const int64_t sl = std::lower_bound(cols.begin(), cols.end(), it->J) - cols.begin();
                const int64_t el = std::lower_bound(cols.begin(), cols.end(), it->J + it->width) - cols.begin();
                const int64_t f0 = std::lower_bound(fcols.begin(), fcols.end(), (int32_t)it->J) - fcols.begin();
                if (el == sl || sl == 0 || f0 >= nF || it->r0 == 0) continue;
                const int64_t Kpad = round_up(el - sl, 64);
                jobs.push_back(Job{sl, el - sl, f0, Kpad});
                maxL = std::max(maxL, round_up(prows(sl) * Kpad, 256));
                maxB = std::max(maxB, round_up(prows(nF - f0) * Kpad, 256));
            }
        const size_t cmap_b = (size_t)round_up(M * 4, 256), fcol_b = (size_t)round_up(std::max<int64_t>(nF, 1) * 4, 256);
        const size_t desc_b = (size_t)round_up((int64_t)(jobs.size() * sizeof(BM)), 256);
        const size_t sched_b = (size_t)round_up((int64_t)jobs.size() * 3 * 8, 256);
        void* pc = nullptr;
        LINAC_TRY(ctx->get(SLOT_PANEL_C, cmap_b + fcol_b + desc_b + sched_b, &pc));
        int32_t* d_cmap = (int32_t*)pc;
        int32_t* d_fcols = (int32_t*)((char*)pc + cmap_b);
        BM* d_desc = (BM*)((char*)pc + cmap_b + fcol_b);
        int64_t* d_sched = (int64_t*)((char*)pc + cmap_b + fcol_b + desc_b);
        LINAC_HIP(hipMemcpyAsync(d_cmap, cmap.data(), M * 4, hipMemcpyHostToDevice, s));
        T* G = nullptr;
        if (nF > 0 && rank > 0) {
            void* pg = nullptr;
            LINAC_TRY(ctx->get(SLOT_OUT, (size_t)(rank * ldg) * sizeof(T), &pg));
            G = (T*)pg;
            LINAC_HIP(hipMemcpyAsync(d_fcols, fcols.data(), nF * 4, hipMemcpyHostToDevice, s));
            k_ech_gather<T><<<dim3((unsigned)ceil_div(nF, 256), (unsigned)std::min<int64_t>(rank, 65535)), 256, 0, s>>>(
                A, ld, ws.piv_rows, d_fcols, rank, nF, G, ldg);
        }
        if (!jobs.empty()) {
            void* pp = nullptr;
            LINAC_TRY(ctx->get(SLOT_PANEL_B, (size_t)NL * (size_t)(maxL + maxB), &pp));
            int8_t* planesL = (int8_t*)pp;
            int8_t* planesB = planesL + NL * maxL;
            std::vector<BM> hd(jobs.size());
            std::vector<int64_t> hs(jobs.size() * 3);
            for (size_t i = 0; i < jobs.size(); ++i) {
                const Job& j = jobs[i];
                BM d{};
                d.A = G + j.f0;
                d.ld = ldg;
                d.N = j.s;
                d.M = nF - j.f0;
                d.f = f;
                d.pd = (double)f.p;
                d.pinv = 1.0 / (double)f.p;
                d.planesL = planesL;
                d.strideL = maxL;
                d.planesB = planesB;
                d.strideB = maxB;
                d.T0 = 0;
                d.nT = nF - j.f0;
                d.Kq = (int32_t)j.Kpad;
                hd[i] = d;
                hs[3 * i] = 0;
                hs[3 * i + 1] = ceil_div(j.s, G_M) * ceil_div(nF - j.f0, TN);
                hs[3 * i + 2] = ceil_div(j.s, G_M);
            }
            LINAC_HIP(hipMemcpyAsync(d_desc, hd.data(), jobs.size() * sizeof(BM), hipMemcpyHostToDevice, s));
            LINAC_HIP(hipMemcpyAsync(d_sched, hs.data(), hs.size() * 8, hipMemcpyHostToDevice, s));
            for (size_t i = 0; i < jobs.size(); ++i) {
                const Job& j = jobs[i];
                const int64_t n = nF - j.f0;
                // L = A[piv_rows[0:s], P_l] (the pivots s .. s + q - 1 give the columns)
                k_planes_cols<LT><<<gs_grid(j.s * j.Kpad, 16384), 256, 0, s>>>(
                    A, ld, ws.piv_rows, ws.piv_cols, j.s, (int)j.q, (int)j.Kpad, j.s, f.p, planesL, maxL);
                // B = G[s:s+q, f0:], rows contiguous
                k_planes_rows<LT><<<dim3((unsigned)std::min<int64_t>(2048, ceil_div(n, 64)), (unsigned)(j.Kpad / 64)),
                                    256, 0, s>>>(G + j.s * ldg + j.f0, ldg, nullptr, 0, (int)j.q, (int)j.Kpad, n, f.p,
                                                 planesB, maxB);
                const TileSched ts{d_sched + 3 * i, d_sched + 3 * i + 2, 1,
                                   (int)tuning().pick(Tuning::GEMM_GROUP, std::max(N, M)), 0, -1};
                const unsigned grid = (unsigned)(ceil_div(hs[3 * i + 1], NUM_XCD) * NUM_XCD);
                if constexpr (u32) {
                    const bool lds_ok = ((((uintptr_t)(G + j.f0)) | (uintptr_t)(ldg * 4)) & 15) == 0 &&
                                        prows(j.s) * j.Kpad < (1ll << 31) && prows(n) * j.Kpad < (1ll << 31);
                    launch_outer32<0>(grid, s, d_desc + i, ts, lds_ok);
                } else {
                    k_gemm_outer64<0><<<grid, G_T, 0, s>>>(d_desc + i, ts);
                }
            }
        }
        constexpr int V = 16 / (int)sizeof(T);
        const int vec = ((ld * (int64_t)sizeof(T)) % 16 == 0 && ((uintptr_t)A & 15) == 0) ? 1 : 0;
        k_ech_write<T><<<dim3((unsigned)std::min<int64_t>(64, ceil_div(ceil_div(M, V), 256)),
                              (unsigned)std::min<int64_t>(N, 65535)), 256, 0, s>>>(A, ld, N, M, G, ldg, d_cmap, rank,
                                                                                    vec);
        LINAC_HIP(hipGetLastError());
        return LINAC_OK;
    }
}

template int echelon_finalize<FieldU32>(uint32_t*, int64_t, int64_t, int64_t, const FieldU32&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, const std::vector<int64_t>&,
                                        const std::vector<int64_t>&, bool, const std::vector<EchelonPanel>&);
template int echelon_finalize<FieldU64>(uint64_t*, int64_t, int64_t, int64_t, const FieldU64&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, const std::vector<int64_t>&,
                                        const std::vector<int64_t>&, bool, const std::vector<EchelonPanel>&);
template int echelon_finalize<FieldGF2>(uint32_t*, int64_t, int64_t, int64_t, const FieldGF2&, hipStream_t,
                                        const ElimWorkspace&, DeviceContext*, const std::vector<int64_t>&,
                                        const std::vector<int64_t>&, bool, const std::vector<EchelonPanel>&);

// ---------------------------------------------------------------------------
// batch objects: matrices join and leave between steps (continuous batching)
// ---------------------------------------------------------------------------

}  // namespace linac

namespace linac {

// Device memory of a batch's items: chunks from hipMalloc, sub-allocated
// best-fit with coalescing free lists, returned to the runtime only when the
// batch is destroyed.  A finished item's memory is reused without hipFree,
// which would synchronise the whole device (and stall the trailing updates
// still running for other items).
class Arena {
  public:
    ~Arena() { release_all(); }
    int alloc(size_t n, void** out) {
        n = (size_t)round_up((int64_t)std::max<size_t>(n, 1), 256);
        Chunk* best = nullptr;
        std::map<size_t, size_t>::iterator bit;
        for (auto& c : chunks)
            for (auto it = c.free.begin(); it != c.free.end(); ++it)
                if (it->second >= n && (!best || it->second < bit->second)) {
                    best = &c;
                    bit = it;
                }
        if (!best) {
            const size_t sz = (size_t)round_up((int64_t)std::max(n, CHUNK), 2 << 20);
            void* p = nullptr;
            const hipError_t e = hipMalloc(&p, sz);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                return fail(LINAC_ERR_NOMEM, "device allocation of %zu bytes failed: %s", sz, hipGetErrorString(e));
            }
            chunks.push_back(Chunk{(char*)p, sz, {}});
            chunks.back().free[0] = sz;
            best = &chunks.back();
            bit = best->free.begin();
        }
        const size_t off = bit->first, len = bit->second;
        best->free.erase(bit);
        if (len > n) best->free[off + n] = len - n;
        *out = best->base + off;
        return LINAC_OK;
    }
    void release(void* p, size_t n) {
        n = (size_t)round_up((int64_t)std::max<size_t>(n, 1), 256);
        for (auto& c : chunks) {
            if ((char*)p < c.base || (char*)p >= c.base + c.size) continue;
            size_t off = (size_t)((char*)p - c.base);
            auto next = c.free.lower_bound(off);
            if (next != c.free.end() && next->first == off + n) {  // merge with the following free block
                n += next->second;
                next = c.free.erase(next);
            }
            if (next != c.free.begin()) {  // and with the preceding one
                auto prev = std::prev(next);
                if (prev->first + prev->second == off) {
                    off = prev->first;
                    n += prev->second;
                    c.free.erase(prev);
                }
            }
            c.free[off] = n;
            return;
        }
    }
    void release_all() {
        for (auto& c : chunks) (void)hipFree(c.base);
        chunks.clear();
    }
    size_t reserved() const {
        size_t t = 0;
        for (auto& c : chunks) t += c.size;
        return t;
    }

  private:
    static constexpr size_t CHUNK = (size_t)1 << 30;
    struct Chunk {
        char* base;
        size_t size;
        std::map<size_t, size_t> free;  // offset -> length
    };
    std::deque<Chunk> chunks;  // (stable addresses)
};

}  // namespace linac

struct linac_batch {
    linac::DeviceContext* ctx = nullptr;
    linac::Engine<linac::LimbsU32>* eng = nullptr;      // 32-bit storage (p <= 4278124287)
    linac::Engine<linac::LimbsU64>* eng64 = nullptr;    // 64-bit storage (2^32 <= p < 2^63), made on first use
    linac::Arena arena;                                  // the items' device memory
    std::vector<linac::Item<linac::LimbsU32>*> items;   // not yet taken
    std::vector<linac::Item<linac::LimbsU32>*> fresh;   // added since the last step (need init)
    std::vector<linac::Item<linac::LimbsU64>*> items64;
    std::vector<linac::Item<linac::LimbsU64>*> fresh64;
    int64_t next_id = 0;  // ids are shared by both widths
    uint64_t cap = 0, in_use = 0;
};

namespace linac {

// The batch's per-width members, selected by storage type.
static std::vector<Item<LimbsU32>*>& items_of(linac_batch* b, LimbsU32) { return b->items; }
static std::vector<Item<LimbsU64>*>& items_of(linac_batch* b, LimbsU64) { return b->items64; }
static std::vector<Item<LimbsU32>*>& fresh_of(linac_batch* b, LimbsU32) { return b->fresh; }
static std::vector<Item<LimbsU64>*>& fresh_of(linac_batch* b, LimbsU64) { return b->fresh64; }
static Engine<LimbsU32>* engine_of(linac_batch* b, LimbsU32) { return b->eng; }
static Engine<LimbsU64>* engine_of(linac_batch* b, LimbsU64) {
    if (!b->eng64) b->eng64 = new Engine<LimbsU64>();
    return b->eng64;
}

template <class LT>
static size_t matrix_bytes(int64_t rows, int64_t cols) {
    return (size_t)round_up(rows * round_up(cols, 64) * (int64_t)sizeof(typename LT::T), 256);
}

template <class LT>
static size_t item_bytes(int64_t rows, int64_t cols) {
    return matrix_bytes<LT>(rows, cols) + workspace_layout<LT>(rows, cols, true).total;
}

template <class LT>
static Item<LT>* find_item(linac_batch* b, int64_t id) {
    for (auto* it : items_of(b, LT{}))
        if (it->id == id) return it;
    return nullptr;
}

// Open batches, so that linac_shutdown (and process exit) can release the
// ones a caller never destroyed, and a handle released that way is refused
// rather than used.
static std::mutex g_batches_mu;
static std::vector<linac_batch*> g_batches;

static bool batch_live(linac_batch* b) {
    std::lock_guard<std::mutex> lk(g_batches_mu);
    return std::find(g_batches.begin(), g_batches.end(), b) != g_batches.end();
}

static void batch_free(linac_batch* b) {
    (void)b->eng->drain();  // no trailing update may still run on an item's memory
    if (b->eng64) (void)b->eng64->drain();
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto* it : b->items) delete it;
    for (auto* it : b->items64) delete it;
    b->arena.release_all();
    delete b->eng;
    delete b->eng64;
    delete b;
}

static void free_context_batches(DeviceContext* ctx) {
    std::vector<linac_batch*> mine;
    {
        std::lock_guard<std::mutex> lk(g_batches_mu);
        for (auto it = g_batches.begin(); it != g_batches.end();) {
            if ((*it)->ctx == ctx) {
                mine.push_back(*it);
                it = g_batches.erase(it);
            } else {
                ++it;
            }
        }
    }
    for (linac_batch* b : mine) batch_free(b);
}

template <class LT>
static int batch_add(linac_batch* b, const typename LT::T* a, int64_t rows, int64_t cols, int64_t ld,
                     const typename LT::F& f, uint32_t flags, int64_t* id_out) {
    using T = typename LT::T;
    const bool in_place = (flags & LINAC_IN_PLACE) != 0;
    if (in_place && !(flags & LINAC_DEVICE_PTR)) return fail(LINAC_ERR_INVALID, "LINAC_IN_PLACE needs LINAC_DEVICE_PTR");
    const size_t need = item_bytes<LT>(rows, cols) - (in_place ? matrix_bytes<LT>(rows, cols) : 0);
    if (b->cap && b->in_use + need > b->cap)
        return fail(LINAC_ERR_NOMEM, "admission: %zu bytes needed, %llu of %llu in use", need,
                    (unsigned long long)b->in_use, (unsigned long long)b->cap);
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    hipStream_t s = b->ctx->stream;
    auto* it = new Item<LT>();
    const int ra = b->arena.alloc(need, &it->arena);
    if (ra != LINAC_OK) {
        delete it;
        return ra;
    }
    it->bytes = need;
    char* base = (char*)it->arena;
    const WorkspaceLayout wl = workspace_layout<LT>(rows, cols, true);
    if (in_place) {
        describe(it->d, const_cast<T*>(a), ld, rows, cols, f);
        carve(it->d, base, wl, rows, cols, true);
        carve_compact(it->dc, base, wl);
    } else {
        const int64_t ldp = round_up(cols, 64);
        T* A = (T*)base;
        hipError_t ce = hipMemcpy2DAsync(A, ldp * sizeof(T), a, ld * sizeof(T), cols * sizeof(T), rows,
                                         (flags & LINAC_DEVICE_PTR) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s);
        if (ce != hipSuccess) {
            b->arena.release(it->arena, it->bytes);
            delete it;
            return fail(LINAC_ERR_DEVICE, "copy into the batch failed: %s", hipGetErrorString(ce));
        }
        describe(it->d, A, ldp, rows, cols, f);
        carve(it->d, base + matrix_bytes<LT>(rows, cols), wl, rows, cols, true);
        carve_compact(it->dc, base + matrix_bytes<LT>(rows, cols), wl);
    }
    it->id = b->next_id++;
    it->forward = tuning().echelon != 0;
    b->in_use += need;
    items_of(b, LT{}).push_back(it);
    fresh_of(b, LT{}).push_back(it);
    *id_out = it->id;
    return LINAC_OK;
}

template <class LT>
static int batch_step(linac_batch* b, hipStream_t s) {
    auto& items = items_of(b, LT{});
    auto& fresh = fresh_of(b, LT{});
    if (items.empty()) return LINAC_OK;
    Engine<LT>* eng = engine_of(b, LT{});
    LINAC_TRY(eng->init_items(fresh, s));
    fresh.clear();
    return eng->step(items, s, nullptr);
}

template <class LT>
static void batch_release_item(linac_batch* b, Item<LT>* it) {
    b->in_use -= it->bytes;
    b->arena.release(it->arena, it->bytes);
    auto& items = items_of(b, LT{});
    items.erase(std::find(items.begin(), items.end(), it));
    delete it;
}

template <class LT>
static int batch_take(linac_batch* b, int64_t id, typename LT::T* out, int64_t ld_out, uint32_t flags,
                      int64_t* pivots_out, int64_t* rank_out) {
    using T = typename LT::T;
    Item<LT>* it = find_item<LT>(b, id);
    if (!it) return fail(LINAC_ERR_INVALID, "unknown or already taken id %lld for %d-bit storage", (long long)id,
                         (int)(8 * sizeof(T)));
    if (!it->finished) return fail(LINAC_ERR_INVALID, "matrix %lld is not finished", (long long)id);
    if (out && ld_out < it->d.M) return fail(LINAC_ERR_INVALID, "ld_out < cols");
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    hipStream_t s = b->ctx->stream;
    LINAC_TRY(engine_of(b, LT{})->drain());  // the item's last trailing-update work has finished
    ElimWorkspace ws{it->d.st, it->d.piv_rows, it->d.piv_cols, it->d.used, std::max<int64_t>(1, std::min(it->d.N, it->d.M))};
    int64_t rank = 0;
    std::vector<int64_t> pc;
    LINAC_TRY(gf_finalize(it->d.A, it->d.ld, it->d.N, it->d.M, it->d.f, s, ws, b->ctx, false, true, &rank, &pc,
                          &it->echelon));
    if (out) {
        LINAC_HIP(hipMemcpy2DAsync(out, ld_out * sizeof(T), it->d.A, it->d.ld * sizeof(T), it->d.M * sizeof(T), it->d.N,
                                   (flags & LINAC_DEVICE_PTR) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
    }
    if (pivots_out)
        for (int64_t k = 0; k < rank; ++k) pivots_out[k] = pc[k];
    if (rank_out) *rank_out = rank;
    batch_release_item(b, it);
    return LINAC_OK;
}

template <class LT>
static int batch_take_many(linac_batch* b, const int64_t* ids, int64_t count, typename LT::T* const* outs,
                           const int64_t* ld_outs, uint32_t flags, int64_t* const* pivots_out, int64_t* ranks_out) {
    using T = typename LT::T;
    using BM = BMat<LT>;
    if (count < 0 || (count > 0 && !ids)) return fail(LINAC_ERR_INVALID, "bad id list");
    if (count == 0) return LINAC_OK;
    std::vector<Item<LT>*> its((size_t)count);
    for (int64_t i = 0; i < count; ++i) {
        its[i] = find_item<LT>(b, ids[i]);
        if (!its[i]) return fail(LINAC_ERR_INVALID, "unknown or already taken id %lld for %d-bit storage",
                                 (long long)ids[i], (int)(8 * sizeof(T)));
        if (!its[i]->finished) return fail(LINAC_ERR_INVALID, "matrix %lld is not finished", (long long)ids[i]);
        for (int64_t j = 0; j < i; ++j)
            if (its[j] == its[i]) return fail(LINAC_ERR_INVALID, "id %lld listed twice", (long long)ids[i]);
        if (outs && outs[i] && (!ld_outs || ld_outs[i] < its[i]->d.M))
            return fail(LINAC_ERR_INVALID, "ld_out < cols for id %lld", (long long)ids[i]);
    }
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    hipStream_t s = b->ctx->stream;
    LINAC_TRY(engine_of(b, LT{})->drain());
    // descriptors, ranks and pivot lists in one scratch block: two round trips in all
    int64_t cap_total = 0;
    for (auto* it : its) cap_total += std::max<int64_t>(1, std::min(it->d.N, it->d.M));
    const size_t desc_b = (size_t)round_up(count * (int64_t)sizeof(BM), 256);
    const size_t offs_b = (size_t)round_up((count + 1) * 8, 256);
    void* scratch = nullptr;
    LINAC_TRY(b->ctx->get(SLOT_AUX, desc_b + offs_b + 2 * (size_t)round_up(cap_total * 8, 256), &scratch));
    BM* d_desc = (BM*)scratch;
    int64_t* d_offs = (int64_t*)((char*)scratch + desc_b);
    int64_t* d_rows = (int64_t*)((char*)scratch + desc_b + offs_b);
    int64_t* d_cols = d_rows + round_up(cap_total * 8, 256) / 8;
    std::vector<BM> hd((size_t)count);
    for (int64_t i = 0; i < count; ++i) hd[i] = its[i]->d;
    LINAC_HIP(hipMemcpyAsync(d_desc, hd.data(), count * sizeof(BM), hipMemcpyHostToDevice, s));
    k_collect_rank<LT><<<(unsigned)count, 64, 0, s>>>(d_desc, d_offs);
    std::vector<int64_t> rk((size_t)count), offs((size_t)count + 1, 0);
    LINAC_HIP(hipMemcpyAsync(rk.data(), d_offs, count * 8, hipMemcpyDeviceToHost, s));
    LINAC_HIP(hipStreamSynchronize(s));
    for (int64_t i = 0; i < count; ++i) offs[i + 1] = offs[i] + rk[i];
    std::vector<int64_t> prow((size_t)std::max<int64_t>(1, offs[count])), pcol(prow.size());
    if (offs[count] > 0) {
        LINAC_HIP(hipMemcpyAsync(d_offs, offs.data(), (count + 1) * 8, hipMemcpyHostToDevice, s));
        k_collect_pivots<LT><<<(unsigned)count, 256, 0, s>>>(d_desc, d_offs, d_rows, d_cols);
        LINAC_HIP(hipMemcpyAsync(prow.data(), d_rows, offs[count] * 8, hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipMemcpyAsync(pcol.data(), d_cols, offs[count] * 8, hipMemcpyDeviceToHost, s));
        LINAC_HIP(hipStreamSynchronize(s));
    }
    for (int64_t i = 0; i < count; ++i) {
        Item<LT>* it = its[i];
        std::vector<int64_t> r(prow.begin() + offs[i], prow.begin() + offs[i + 1]);
        std::vector<int64_t> c(pcol.begin() + offs[i], pcol.begin() + offs[i + 1]);
        const bool resorted = sort_pivots_by_column(r, c);
        ElimWorkspace ws{it->d.st, it->d.piv_rows, it->d.piv_cols, it->d.used,
                         std::max<int64_t>(1, std::min(it->d.N, it->d.M))};
        if (it->echelon.empty())
            LINAC_TRY(gf_finalize_apply(it->d.A, it->d.ld, it->d.N, it->d.M, it->d.f, s, ws, b->ctx, false, r, c,
                                        resorted));
        else
            LINAC_TRY(echelon_finalize(it->d.A, it->d.ld, it->d.N, it->d.M, it->d.f, s, ws, b->ctx, r, c, resorted,
                                       it->echelon));
        if (outs && outs[i])
            LINAC_HIP(hipMemcpy2DAsync(outs[i], ld_outs[i] * sizeof(T), it->d.A, it->d.ld * sizeof(T),
                                       it->d.M * sizeof(T), it->d.N,
                                       (flags & LINAC_DEVICE_PTR) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                       s));
        if (pivots_out && pivots_out[i])
            for (int64_t k = 0; k < (int64_t)c.size(); ++k) pivots_out[i][k] = c[k];
        if (ranks_out) ranks_out[i] = rk[i];
    }
    LINAC_HIP(hipStreamSynchronize(s));
    for (auto* it : its) batch_release_item(b, it);
    return LINAC_OK;
}

static int batch_check_add(linac_batch* b, const void* a, int64_t rows, int64_t cols, int64_t ld, int64_t* id_out) {
    if (!b || !id_out) return fail(LINAC_ERR_INVALID, "batch or id_out is NULL");
    if (!batch_live(b)) return fail(LINAC_ERR_INVALID, "batch handle was released (linac_shutdown)");
    if (rows <= 0 || cols <= 0) return fail(LINAC_ERR_INVALID, "empty matrix");
    if (ld < cols) return fail(LINAC_ERR_INVALID, "ld < cols");
    if (!a) return fail(LINAC_ERR_INVALID, "matrix pointer is NULL");
    return LINAC_OK;
}

}  // namespace linac

using namespace linac;

extern "C" {

int linac_batch_footprint_u32(int64_t rows, int64_t cols, uint64_t* bytes_out) {
    if (rows < 0 || cols < 0 || !bytes_out) return fail(LINAC_ERR_INVALID, "bad arguments");
    *bytes_out = (rows == 0 || cols == 0) ? 0 : (uint64_t)item_bytes<LimbsU32>(rows, cols);
    return LINAC_OK;
}

int linac_batch_footprint_u64(int64_t rows, int64_t cols, uint64_t* bytes_out) {
    if (rows < 0 || cols < 0 || !bytes_out) return fail(LINAC_ERR_INVALID, "bad arguments");
    *bytes_out = (rows == 0 || cols == 0) ? 0 : (uint64_t)item_bytes<LimbsU64>(rows, cols);
    return LINAC_OK;
}

int linac_batch_create(linac_batch** out, uint64_t mem_cap_bytes) {
    if (!out) return fail(LINAC_ERR_INVALID, "out is NULL");
    DeviceContext* ctx = nullptr;
    LINAC_TRY(current_context(&ctx));
    auto* b = new linac_batch();
    b->ctx = ctx;
    b->eng = new Engine<LimbsU32>();
    b->cap = mem_cap_bytes;
    add_shutdown_hook(free_context_batches);
    {
        std::lock_guard<std::mutex> lk(g_batches_mu);
        g_batches.push_back(b);
    }
    *out = b;
    return LINAC_OK;
}

int linac_batch_destroy(linac_batch* b) {
    if (!b) return LINAC_OK;
    {
        std::lock_guard<std::mutex> lk(g_batches_mu);
        auto it = std::find(g_batches.begin(), g_batches.end(), b);
        if (it == g_batches.end()) return LINAC_OK;  // already released by linac_shutdown
        g_batches.erase(it);
    }
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    batch_free(b);
    return LINAC_OK;
}

int linac_batch_memory(linac_batch* b, uint64_t* in_use_out, uint64_t* cap_out) {
    if (!b) return fail(LINAC_ERR_INVALID, "batch is NULL");
    if (in_use_out) *in_use_out = b->in_use;
    if (cap_out) *cap_out = b->cap;
    return LINAC_OK;
}

int linac_batch_add_u32(linac_batch* b, const uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p,
                        uint32_t flags, int64_t* id_out) {
    LINAC_TRY(batch_check_add(b, a, rows, cols, ld, id_out));
    if (!is_prime_u64(p) || p == 2) return fail(LINAC_ERR_INVALID, "field characteristic %u is not an odd prime", p);
    if (p > LIMB_P_MAX) return fail(LINAC_ERR_INVALID, "p = %u exceeds the batched path's limit %llu", p,
                                    (unsigned long long)LIMB_P_MAX);
    return batch_add<LimbsU32>(b, a, rows, cols, ld, make_field_u32(p), flags, id_out);
}

int linac_batch_add_u64(linac_batch* b, const uint64_t* a, int64_t rows, int64_t cols, int64_t ld, uint64_t p,
                        uint32_t flags, int64_t* id_out) {
    LINAC_TRY(batch_check_add(b, a, rows, cols, ld, id_out));
    if (p < LIMB64_P_MIN || p >= LIMB64_P_MAX || !is_prime_u64(p))
        return fail(LINAC_ERR_INVALID, "p = %llu is not a prime in [2^32, 2^63) (smaller primes: linac_batch_add_u32)",
                    (unsigned long long)p);
    return batch_add<LimbsU64>(b, a, rows, cols, ld, make_field_u64(p), flags, id_out);
}

int linac_batch_step(linac_batch* b, int64_t* unfinished_out) {
    if (!b) return fail(LINAC_ERR_INVALID, "batch is NULL");
    if (!batch_live(b)) return fail(LINAC_ERR_INVALID, "batch handle was released (linac_shutdown)");
    std::lock_guard<std::recursive_mutex> lk(b->ctx->mu);
    hipStream_t s = b->ctx->stream;
    LINAC_TRY(batch_step<LimbsU32>(b, s));
    LINAC_TRY(batch_step<LimbsU64>(b, s));
    LINAC_HIP(hipStreamSynchronize(s));
    int64_t n = 0;
    for (auto* it : b->items) n += it->finished ? 0 : 1;
    for (auto* it : b->items64) n += it->finished ? 0 : 1;
    if (unfinished_out) *unfinished_out = n;
    return LINAC_OK;
}

int linac_batch_run(linac_batch* b) {
    int64_t n = 1;
    while (n > 0) LINAC_TRY(linac_batch_step(b, &n));
    return LINAC_OK;
}

int linac_batch_status(linac_batch* b, int64_t id) {
    if (!b) return fail(LINAC_ERR_INVALID, "batch is NULL");
    if (Item<LimbsU32>* it = find_item<LimbsU32>(b, id)) return it->finished ? 1 : 0;
    if (Item<LimbsU64>* it = find_item<LimbsU64>(b, id)) return it->finished ? 1 : 0;
    return fail(LINAC_ERR_INVALID, "unknown or already taken id %lld", (long long)id);
}

int linac_batch_take_u32(linac_batch* b, int64_t id, uint32_t* out, int64_t ld_out, uint32_t flags,
                         int64_t* pivots_out, int64_t* rank_out) {
    if (!b) return fail(LINAC_ERR_INVALID, "batch is NULL");
    return batch_take<LimbsU32>(b, id, out, ld_out, flags, pivots_out, rank_out);
}

int linac_batch_take_u64(linac_batch* b, int64_t id, uint64_t* out, int64_t ld_out, uint32_t flags,
                         int64_t* pivots_out, int64_t* rank_out) {
    if (!b) return fail(LINAC_ERR_INVALID, "batch is NULL");
    return batch_take<LimbsU64>(b, id, out, ld_out, flags, pivots_out, rank_out);
}

int linac_batch_take_many_u32(linac_batch* b, const int64_t* ids, int64_t count, uint32_t* const* outs,
                              const int64_t* ld_outs, uint32_t flags, int64_t* const* pivots_out,
                              int64_t* ranks_out) {
    if (!b) return fail(LINAC_ERR_INVALID, "batch is NULL");
    return batch_take_many<LimbsU32>(b, ids, count, outs, ld_outs, flags, pivots_out, ranks_out);
}

int linac_batch_take_many_u64(linac_batch* b, const int64_t* ids, int64_t count, uint64_t* const* outs,
                              const int64_t* ld_outs, uint32_t flags, int64_t* const* pivots_out,
                              int64_t* ranks_out) {
    if (!b) return fail(LINAC_ERR_INVALID, "batch is NULL");
    return batch_take_many<LimbsU64>(b, ids, count, outs, ld_outs, flags, pivots_out, ranks_out);
}

int linac_gemm_sub_u32(uint32_t* C, int64_t ldc, const uint32_t* A, int64_t lda, int64_t a_ncols, const int64_t* a_cols,
                       const uint32_t* B, int64_t ldb, int64_t b_nrows, const int64_t* b_rows, int64_t m, int64_t n,
                       int64_t k, uint32_t p, uint32_t flags) {
    return gemm_sub_u32(C, ldc, A, lda, a_ncols, a_cols, B, ldb, b_nrows, b_rows, m, n, k, p, flags);
}

int linac_rref_batch_u32(uint32_t* const* mats, const int64_t* rows, const int64_t* cols, const int64_t* lds,
                         const uint32_t* primes, int64_t count, uint32_t flags, int64_t* const* pivots_out,
                         int64_t* ranks_out) {
    if (count < 0 || (count > 0 && (!mats || !rows || !cols || !lds || !primes)))
        return fail(LINAC_ERR_INVALID, "bad batch arguments");
    linac_batch* b = nullptr;
    LINAC_TRY(linac_batch_create(&b, 0));
    std::vector<int64_t> ids(count, -1);
    int rc = LINAC_OK;
    for (int64_t i = 0; i < count && rc == LINAC_OK; ++i) {
        if (rows[i] == 0 || cols[i] == 0) {
            if (ranks_out) ranks_out[i] = 0;
            continue;
        }
        rc = linac_batch_add_u32(b, mats[i], rows[i], cols[i], lds[i], primes[i], flags, &ids[i]);
    }
    if (rc == LINAC_OK) rc = linac_batch_run(b);
    if (rc == LINAC_OK) {
        std::vector<int64_t> tid, tld, trk;
        std::vector<uint32_t*> tout;
        std::vector<int64_t*> tpiv;
        std::vector<int64_t> where;
        for (int64_t i = 0; i < count; ++i) {
            if (ids[i] < 0) continue;
            tid.push_back(ids[i]);
            tout.push_back(mats[i]);
            tld.push_back(lds[i]);
            tpiv.push_back(pivots_out ? pivots_out[i] : nullptr);
            where.push_back(i);
        }
        trk.resize(tid.size());
        rc = linac_batch_take_many_u32(b, tid.data(), (int64_t)tid.size(), tout.data(), tld.data(), flags,
                                       tpiv.data(), trk.data());
        if (rc == LINAC_OK && ranks_out)
            for (size_t k = 0; k < where.size(); ++k) ranks_out[where[k]] = trk[k];
    }
    const int rc2 = linac_batch_destroy(b);
    return rc != LINAC_OK ? rc : rc2;
}

}  // extern "C"
