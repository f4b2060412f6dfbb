// Declarations shared by the GF(p) elimination methods.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "field.hpp"
#include "runtime.hpp"

namespace linac {

// 16-byte vector of field elements: one global_load/store_dwordx4 per access.
template <class T>
struct alignas(16) Vec16 {
    T v[16 / sizeof(T)];
};

// Return value of a method that does not apply to a problem (distinct from
// the negative error codes): the caller falls back to another method.
constexpr int NOT_APPLICABLE = 1;

// An outer panel [J, J + width) of the blocked elimination whose updates
// skipped the rows above r0 (all of them pivot rows of earlier panels): those
// rows still hold their entries in the panel's pivot columns, and the back
// substitution of echelon_finalize removes them.
struct EchelonPanel {
    int64_t J, width, r0;
};

// Blocked right-looking elimination (rref_blocked.hip).  Runs the whole
// elimination on A (state in `ws`), leaving every pivot row normalised.
// With `echelon` null every row takes every panel's update (Gauss-Jordan:
// the pivot rows are final).  Otherwise a pivot row leaves the updates once
// every row above the next panel's first unused row is a pivot row (forward
// elimination to echelon form, about 2/3 of the multiply-adds for a square
// matrix), and the panels that skipped rows are appended to *echelon for
// echelon_finalize.  Returns LINAC_OK, NOT_APPLICABLE, or a negative error.
template <class F>
int blocked_dispatch(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                     const ElimWorkspace& ws, DeviceContext* ctx, bool forced, Profiler* prof,
                     std::vector<EchelonPanel>* echelon = nullptr);

// Shared finalisation (rref_gf.hip): reads the pivot list, optionally sorts it
// by column and normalises, and gathers the pivot rows to rows 0..rank-1.
// With `echelon` panels that skipped rows, the result is completed by
// echelon_finalize instead of the gather.
template <class F>
int gf_finalize(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                const ElimWorkspace& ws, DeviceContext* ctx, bool normalize, bool sort_by_column,
                int64_t* rank_host, std::vector<int64_t>* piv_cols_host,
                const std::vector<EchelonPanel>* echelon = nullptr);
// The gather alone, enqueued without synchronisation, for a pivot list the
// caller has already read (and sorted: `resorted` rewrites the device copy).
template <class F>
int gf_finalize_apply(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                      const ElimWorkspace& ws, DeviceContext* ctx, bool normalize, const std::vector<int64_t>& rows,
                      const std::vector<int64_t>& cols, bool resorted);
bool sort_pivots_by_column(std::vector<int64_t>& rows, std::vector<int64_t>& cols);
// Completes an echelon-form elimination (rref_blocked.hip), enqueued on s
// without synchronisation: the entries of the pivot rows in the non-pivot
// columns are gathered (rows in result order), the block back substitution
// runs panel by panel from the last (exact limb GEMMs restricted to those
// columns), and A is rewritten as the RREF -- identity in the pivot columns,
// the gathered values elsewhere, zero rows below the rank.  rows/cols: the
// pivot list sorted by column (`resorted` rewrites the device copy).  Panels
// list empty: the plain gather (gf_finalize_apply) applies.
template <class F>
int echelon_finalize(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                     const ElimWorkspace& ws, DeviceContext* ctx, const std::vector<int64_t>& rows,
                     const std::vector<int64_t>& cols, bool resorted, const std::vector<EchelonPanel>& panels);

// The plain method on a device matrix with all of its buffers carved from
// `scratch` (rref_gf.hip); noalloc_scratch_bytes is what it needs.
size_t noalloc_scratch_bytes(int64_t rows, int64_t cols, int64_t ld, size_t elem);
int rref_u32_noalloc(uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p, uint32_t flags, hipStream_t s,
                     DeviceContext* ctx, uint8_t* scratch, int64_t* pivots_out, int64_t* rank_out);

}  // namespace linac
