// Internal runtime of the linac_amd library: per-device context, error
// reporting, cached device scratch, and the bookkeeping that every
// elimination method shares.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "linac_amd.h"

namespace linac {

// Records a message for linac_last_error() and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define LINAC_HIP(expr)                                                                    \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return ::linac::fail(LINAC_ERR_DEVICE, "%s: %s (%s:%d)", #expr,                \
                                 hipGetErrorString(_e), __FILE__, __LINE__);               \
    } while (0)

#define LINAC_TRY(expr)            \
    do {                           \
        int _rc = (expr);          \
        if (_rc != LINAC_OK) return _rc; \
    } while (0)

// Grow-only device buffers owned by a context, reused across calls so that a
// steady stream of solves does not hit hipMalloc.
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};

enum ScratchSlot {
    SLOT_STATE = 0,   // ElimState + pivot lists + row flags
    SLOT_MATRIX,      // device copy of a host matrix
    SLOT_OUT,         // permutation target of the final gather
    SLOT_PANEL_A,     // blocked method: panel / limb planes
    SLOT_PANEL_B,
    SLOT_PANEL_C,
    SLOT_AUX,
    SLOT_AUX2,        // float refinement: kept panel factors
    SLOT_JOB_A,       // resumable float job: sweep workspace
    SLOT_JOB_B,       // resumable float job: state, pivot lists, row flags
    SLOT_COUNT
};

struct DeviceContext {
    int device = -1;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    Scratch scratch[SLOT_COUNT];
    int64_t* pinned = nullptr;    // small pinned host buffer for status reads
    hipEvent_t probe = nullptr;   // reusable timing-free event (rank polls of the plain sweep)
    void* engines[2] = {nullptr, nullptr};  // cached blocked-elimination engines (32- and 64-bit storage)
    void* float_job = nullptr;              // the device's resumable float job, if one is open
    int get(ScratchSlot slot, size_t bytes, void** out);
    int release();
};

// Called with the context's lock held, before its stream is destroyed, by
// linac_shutdown and at process exit: frees what other parts of the library
// cache in a context (engines, batches).
typedef void (*ShutdownHook)(DeviceContext* ctx);
void add_shutdown_hook(ShutdownHook fn);

// Context of the current device (created on first use).  Fails with
// LINAC_ERR_FORKED in a child process forked after initialisation and with
// LINAC_ERR_DEVICE when no GPU is usable.
int current_context(DeviceContext** out);
// The current device's context if it has been created (by linac_init or an
// earlier call), without creating one; LINAC_ERR_INVALID otherwise.
int existing_context(DeviceContext** out);

// No-allocation entry point (noalloc.hip): resolves the current device's
// address of the library's static scratch array.  Called by linac_init.
int noalloc_setup(int device);

// Shared state of one elimination (lives in device memory).
struct ElimState {
    int64_t prow;          // pivot row of the current column, -1 if none
    int64_t rank;          // pivots found so far
    int64_t first_unused;  // every row below this index is a pivot row
    uint64_t factor;       // field-specific multiplier derived from the head
    int64_t aux[4];
};

struct ElimWorkspace {
    ElimState* st;
    int64_t* piv_rows;  // [min(N, M)] row of the k-th pivot
    int64_t* piv_cols;  // [min(N, M)] column of the k-th pivot
    uint8_t* used;      // [N] 1 once a row has been a pivot row
    int64_t cap;        // min(N, M)
};

int make_workspace(DeviceContext* ctx, int64_t N, int64_t M, ElimWorkspace* ws);

bool is_prime_u64(uint64_t n);

// Event-based timing of one call (LINAC_PROFILE).  Samples are launches of
// the method's dominant kernel, each tagged with its algorithmic bytes and
// modular multiply-adds so that achieved rates need no model of the kernel.
struct Profiler {
    bool on = false;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    struct Sample {
        hipEvent_t a, b;
        double bytes, macs;
        int64_t tag;  // column (plain) or panel start (blocked)
    };
    std::vector<Sample> samples;
    double result[6] = {0, 0, 0, 0, 0, 0};  // see linac_profile_last
    ~Profiler();
    void clear();                                    // destroys every event
    int begin(hipStream_t s);                        // resets and records t0
    int sample_begin(hipStream_t s, int64_t tag, double bytes, double macs);
    int sample_end(hipStream_t s);
    int finish(hipStream_t s);                       // records t1, syncs, sums
};
Profiler& thread_profiler();

// Process-wide tuning knobs (linac_tune_set); read at launch time.
struct Tuning {
    int64_t gf_blocked_min_rows = 1024, gf_blocked_min_cols = 512;
    int64_t float_blocked_min_rows = 256, float_blocked_min_cols = 128;
    int64_t gemm_group = 8;
    int64_t lookahead = 1;
    int64_t timeslice = 0;
    int64_t gemm_lds = 1;
    int64_t fast_panel = 2;
    int64_t echelon = 1;  // blocked GF(p): forward elimination plus back substitution (0: Gauss-Jordan)
    int64_t inner_compact = 0;  // 32-bit inner factorisation in the compact form (32 slots a row)
    int64_t inner_threads = 0;  // inner factorisation: threads per workgroup (256, 512, 1024; 0: 1024 for 32-bit, 512 for 64-bit)
    int64_t fpanel_rows = 16;   // float64 panel kernel: rows per wave (16, or 32 for half the workgroups)
    int64_t tile64 = 1;         // 64-bit trailing GEMM tile: 64 x 32 (1) or 64 x 64 (2)
    int64_t u_passes = 0;       // column pairs per U workgroup (1, 2, 4, 8; 0: 1 below 8 matrices, 4 from 8 on)
    int64_t inner_split = 16;   // workgroups per row stripe of the inner GEMM at most (0: 8 for 64-bit, 4 for 32-bit)
    // Per size class overrides of the knobs above that shape the GEMMs (0: the
    // global value applies).  The class of a step or call is that of its
    // largest matrix, n = max(rows, cols): small n < 3072, medium n < 12288,
    // large otherwise.  Keys "<knob>.s", "<knob>.m", "<knob>.l".
    enum Sized { GEMM_GROUP, INNER_SPLIT, TILE64, FPANEL_ROWS, NSIZED };
    // (float64 panels: 32 rows per wave from the medium class on, measured
    // 5-8% faster at N = 6000-8000, 10-15% slower at N <= 2000)
    int64_t by_class[NSIZED][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 32, 32}};
    int64_t pick(Sized k, int64_t n) const {
        const int c = n < 3072 ? 0 : n < 12288 ? 1 : 2;
        const int64_t v = by_class[k][c];
        if (v != 0) return v;
        return k == GEMM_GROUP ? gemm_group : k == INNER_SPLIT ? inner_split : k == TILE64 ? tile64 : fpanel_rows;
    }
};
Tuning& tuning();

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
__host__ __device__ inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

}  // namespace linac
