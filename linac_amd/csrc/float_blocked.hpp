// The blocked float64 / complex128 sweep as a resumable object: set-up,
// panels in any number of run() calls, finish() (error check and the pivot
// counts).  rref_float_blocked.hip implements it; the single-call path runs
// it to the end at once, float jobs (rref_float.hip) a few panels per call.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "float_ops.hpp"
#include "float_refine.hpp"
#include "runtime.hpp"

namespace linac {

// One workgroup's candidate for a column.  Both words carry the column's
// stamp in bit 63 (|x| >= 0 leaves the sign bit free; rows are < 2^62), so a
// reader that sees the stamp in both words has a consistent pair without any
// ordering between the two stores.
struct FSlot {
    uint64_t val;  // bits of |x|
    uint64_t row;  // row index, or ROW_NONE
};

template <class T>
struct FPanelArgs {
    T* A;
    int64_t ld, N, M;
    int64_t J;        // first column of the panel
    int w;            // its width (<= FB)
    int panel;        // panel index
    double tol;
    T* ZL;            // [N][FB]  -P[r, pc] (multipliers), zero rows for this panel's pivots
    T* ZQ;            // [FB][FB] Z rows of this panel's pivots, by slot
    int64_t* qrows;   // [FB]     pivot rows of this panel, by slot
    int32_t* kpanel;  // pivots of this panel
    int32_t* pepoch;  // [N] panel in which the row became a pivot (-1: none)
    uint8_t* used;    // [N]
    int64_t* piv_rows;
    int64_t* piv_cols;
    int64_t* rank;
    FSlot* slots;     // [2][G], alternating by column
    T* crow;          // [2][G][2 * FB] each workgroup's candidate row (panel | Z)
    int* err;
    int64_t spin_limit;
    unsigned long long* tphase;  // phase timers (workgroup 0, thread 0)
};

template <class T>
struct FloatSweep {
    FPanelArgs<T> P{};
    int32_t* kpanel = nullptr;    // [npanels] pivots of each panel
    double* BQ = nullptr;         // right operand of the trailing update
    const void* kfn = nullptr;    // the cooperative panel kernel (rows per wave chosen at set-up)
    int64_t G = 0;                // its workgroups
    unsigned lds = 0;             // its dynamic LDS
    int64_t npanels = 0, next = 0;
    DeviceContext* ctx = nullptr;
    FloatFactors* keep = nullptr;
    bool f32_update = false;      // the trailing updates on FP32 matrix cores (the mixed method)
    // Returns 1 when the cooperative panel grid cannot be resident on this
    // device (the caller then takes the plain method).  Workspace: `ws` (at
    // least workspace_bytes(N, M)) when given, else the context's `slot`.
    int setup(T* A, int64_t ld, int64_t N, int64_t M, double tol, hipStream_t s, int64_t* rank, uint8_t* used,
              int64_t* piv_rows, int64_t* piv_cols, DeviceContext* ctx, ScratchSlot slot, FloatFactors* keep,
              void* ws = nullptr);
    static size_t workspace_bytes(int64_t N, int64_t M);  // upper bound over the panel kernel's shapes
    int run(hipStream_t s, int64_t max_panels, Profiler* prof);  // max_panels < 0: to the end
    int finish(hipStream_t s, Profiler* prof);
    bool done() const { return next >= npanels; }
};

}  // namespace linac
