// Residual-based refinement of the float RREF's solution columns.
//
// The blocked float sweep (rref_float_blocked.hip) transforms the row-scaled
// matrix S A panel by panel: for panel i with pivot rows Q_i (slot order),
//   pivot rows:   X[Q_i, :] <- ZQ_i X[Q_i, :]
//   other rows:   X[r, :]   <- X[r, :] + ZL_i[r, :] (ZQ_i X[Q_i, :])
// (ZL_i = -(the panel's values in its pivot columns), ZQ_i = inverse of the
// pivot block).  With the factors of every panel kept, the same
// transformation can be replayed on extra columns.  For the non-pivot
// ("solution") columns NP of the result, X = R[Q, NP] should satisfy
// A[:, P] X = A[:, NP]; the residual E = A[:, NP] - A[:, P] X is formed in
// double-double arithmetic from a copy of the original matrix, scaled by the
// row scales, replayed, and its pivot-row part added to X.  Each pass reduces
// the solution error by about cond(A) * 2^-53 (plus the residual's own
// double-double rounding), so a pass or two reach the accuracy of the exact
// solution rounded to float64 for cond(A) well below 2^53.
#pragma once
#include <stdint.h>

#include <vector>

#include "runtime.hpp"

namespace linac {

// Per-panel factors of one blocked float sweep (device arrays; host copy of
// the pivot counts).  Element type T (double or CplxF).
struct FloatFactors {
    void* ZL = nullptr;          // [npanels][N][FB]
    void* ZQ = nullptr;          // [npanels][FB][FB]
    int64_t* qrows = nullptr;    // [npanels][FB]
    const int32_t* pepoch = nullptr;  // [N] panel in which a row became a pivot (-1: none)
    std::vector<int32_t> kp;     // pivots per panel (host)
    int64_t npanels = 0;
};

// Statistics of this thread's last refined call (linac_refine_last).
struct RefineStats {
    double iterations = 0;
    double first_correction = 0;  // max |correction| / max |X| of the first pass
    double last_correction = 0;   // ... of the last pass
    double solution_columns = 0;
    double fallback = 0;          // 1: a mixed call's FP32 run did not refine to FP64 accuracy and was redone in FP64
};
RefineStats& thread_refine_stats();

// Refines the non-pivot columns of the blocked float result A (before the
// final gather: pivot rows in place, normalised).  Aorig: the original
// matrix (unscaled), same shape, row stride ldo.  piv_rows / piv_cols: the
// device pivot lists (rank entries, in column order); pc_host: the pivot
// columns on the host.  Runs up to `passes` passes on stream s; with
// stop_below > 0 it stops after a pass whose max |correction| / max |X| is at
// most stop_below, or that corrected more than half as much as the pass before
// (the iteration is not converging).
template <class T>
int float_refine(T* A, int64_t ld, const T* Aorig, int64_t ldo, int64_t N, int64_t M, const int64_t* piv_rows,
                 const int64_t* piv_cols, const std::vector<int64_t>& pc_host, const FloatFactors& fac, int passes,
                 hipStream_t s, DeviceContext* ctx, double stop_below = 0.0);

}  // namespace linac
