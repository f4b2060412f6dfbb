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

// Blocked right-looking Gauss-Jordan (rref_blocked.hip).  Runs the whole
// elimination on A (state in `ws`), leaving every pivot row normalised.
// Returns LINAC_OK, NOT_APPLICABLE, or a negative error.
template <class F>
int blocked_dispatch(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                     const ElimWorkspace& ws, DeviceContext* ctx, bool forced, Profiler* prof);

// Shared finalisation (rref_gf.hip): reads the pivot list, optionally sorts it
// by column and normalises, and gathers the pivot rows to rows 0..rank-1.
template <class F>
int gf_finalize(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                const ElimWorkspace& ws, DeviceContext* ctx, bool normalize, bool sort_by_column,
                int64_t* rank_host, std::vector<int64_t>* piv_cols_host);
// The gather alone, enqueued without synchronisation, for a pivot list the
// caller has already read (and sorted: `resorted` rewrites the device copy).
template <class F>
int gf_finalize_apply(typename F::T* A, int64_t ld, int64_t N, int64_t M, const F& f, hipStream_t s,
                      const ElimWorkspace& ws, DeviceContext* ctx, bool normalize, const std::vector<int64_t>& rows,
                      const std::vector<int64_t>& cols, bool resorted);
bool sort_pivots_by_column(std::vector<int64_t>& rows, std::vector<int64_t>& cols);

// The plain method on a device matrix with all of its buffers carved from
// `scratch` (rref_gf.hip); noalloc_scratch_bytes is what it needs.
size_t noalloc_scratch_bytes(int64_t rows, int64_t cols, int64_t ld, size_t elem);
int rref_u32_noalloc(uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p, uint32_t flags, hipStream_t s,
                     DeviceContext* ctx, uint8_t* scratch, int64_t* pivots_out, int64_t* rank_out);

}  // namespace linac
