/*
 * linac_amd -- C ABI of the MI355X (gfx950) dense row-reduction library.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns an
 * int status (LINAC_OK or a negative LINAC_ERR_*) and leaves a message for
 * linac_last_error().  Matrices are row-major with a leading dimension `ld`
 * (elements between consecutive rows, ld >= cols).
 *
 * Pointers are host memory unless LINAC_DEVICE_PTR is set in `flags`; with a
 * host pointer the library copies to HBM, runs, and copies back in place.  With
 * a device pointer the rows must start on 16-byte boundaries (ld * element size
 * a multiple of 16 and a 16-byte aligned base).  `stream` may be NULL (the
 * library's own stream for the current device).  All calls block until their
 * results are complete.
 *
 * Reference interfaces these replace (linac 1.0.1):
 *   linac_rref_*         cuda_row_reduce            linac/pycuda_row_reduce.py:18-151
 *                        (the JIT+launch glue of    linac/pycuda_tools.py:17-49 and the
 *                         kernels of                linac/row_reduce.cu:117-430)
 *   linac_load_matrix_*  cuda_load_matrix           linac/pycuda_matrix_loader.py:28-124
 *                        (kernel                    linac/matrix_loader.cu:88-112)
 *   linac_available      CUDA_FOUND                 linac/_optional.py:11-23
 */
#ifndef LINAC_AMD_H
#define LINAC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped to Python exceptions by the host package) ---- */
#define LINAC_OK                   0
#define LINAC_ERR_INVALID         -1  /* bad argument: ValueError                   */
#define LINAC_ERR_NOMEM           -2  /* allocation or admission: MemoryError        */
#define LINAC_ERR_DEVICE          -3  /* HIP runtime / no device: RuntimeError       */
#define LINAC_ERR_NOT_INVERTIBLE  -4  /* head not invertible mod p: ZeroDivisionError */
#define LINAC_ERR_FORKED          -5  /* used after fork() in a child: RuntimeError   */

/* ---- flags ---- */
#define LINAC_DEVICE_PTR      0x1u   /* matrix / operand pointers are device memory */
#define LINAC_METHOD_MASK     0xF0u
#define LINAC_METHOD_AUTO     0x00u  /* library picks (blocked when it applies)     */
#define LINAC_METHOD_PLAIN    0x10u  /* unblocked: one fused step per column        */
#define LINAC_METHOD_BLOCKED  0x20u  /* panel + exact mod-p MFMA trailing update    */
#define LINAC_METHOD_MIXED    0x30u  /* float64 only: FP32 trailing updates + refinement */
#define LINAC_NO_FINALIZE     0x100u /* leave rows unpermuted (pivots_out says where)*/
#define LINAC_PROFILE         0x200u /* time the call with HIP events (linac_profile_last) */
#define LINAC_IN_PLACE        0x400u /* batch: work on the caller's device matrix, no copy */
#define LINAC_REFINE          0x800u /* float: refine the solution columns (blocked method) */

/* ---- library ---- */
const char* linac_last_error(void);
int linac_version(void);
/* Number of visible GPUs (0 when none); never fails on a machine without one. */
int linac_device_count(int* count);
/* 1 when a usable GPU is present, else 0: the meaning of CUDA_FOUND
 * (linac/_optional.py:11-23). */
int linac_available(void);
/* Create the per-device context (stream, workspace).  Optional: every entry
 * point initialises lazily on the current device. */
int linac_init(int device);
/* Synchronise and release cached device memory of the current device. */
int linac_release_cache(void);
/* Synchronise every device this process used and release everything the
 * library holds there: cached engines, open batches (their handles become
 * invalid), scratch, the pinned status page, events and streams.  Later calls
 * start afresh.  Also runs by itself at process exit (registered with atexit
 * once the HIP runtime is up, so it runs before the runtime's own teardown).
 * No counterpart in the reference, which leaves its context to PyCUDA's
 * autoinit (linac/_optional.py:15-17). */
int linac_shutdown(void);
/* 1 if n is prime (deterministic Miller-Rabin), else 0.  Host only. */
int linac_is_prime(uint64_t n);
/* HIP-event timings of this thread's last call made with LINAC_PROFILE, all
 * recorded on the call's stream:
 *   out[0] device ms of the elimination (first to last kernel),
 *   out[1] sampled launches of the dominant kernel (plain: row update;
 *          blocked: trailing mod-p GEMM),
 *   out[2] their summed duration in ms,
 *   out[3] their algorithmic HBM bytes, out[4] their algorithmic modular MACs,
 *   out[5] the method that ran (1 plain, 2 blocked).
 * Writes min(n, 6) values and returns how many were written. */
int linac_profile_last(double* out, int n);
/* Run-time tuning knobs, process-wide (no counterpart in the reference, whose
 * kernels are specialised by JIT per shape).  linac_amd.autotune measures and
 * sets them.  Keys and defaults:
 *   "gf_blocked_min_rows" 1024, "gf_blocked_min_cols" 512: LINAC_METHOD_AUTO
 *       picks the blocked GF(p) method from this size on;
 *   "float_blocked_min_rows" 256, "float_blocked_min_cols" 128: the same for
 *       float64 / complex128;
 *   "gemm_group" 8: tile rows walked together in the GEMM tile order (a right
 *       tile is reused that many times while the left tiles stay in L2), 1..64;
 *   "lookahead" 1: the trailing update of a panel overlaps the next panel's
 *       inner steps (0: one after the other);
 *   "timeslice" 0: with look-ahead, each part of the trailing update beyond
 *       the next panel runs as one launch beside the next inner steps (1: in
 *       slices beside each inner step's factorisation, the inner step's GEMM
 *       waiting for its slice);
 *   "gemm_lds" 1: the 32-bit trailing GEMM brings its operand stages and its
 *       C tile into LDS with direct-to-LDS loads where every matrix's rows are
 *       16-byte aligned (0: register-staged copies everywhere);
 *   "fast_panel" 2: full-rank panels through their candidate block with the
 *       inner-step kernels (1: one workgroup per block; 0: off);
 *   "echelon" 1: blocked GF(p) by forward elimination to echelon form and a
 *       block back substitution (0: Gauss-Jordan form throughout); results
 *       are the same (the RREF is unique);
 *   "inner_compact" 0: the 32-bit inner factorisation in the compact form;
 *   "inner_threads" 0: threads of the inner factorisation's workgroup, 256,
 *       512 or 1024 (fewer waves, each with more rows; 0: 1024 for 32-bit,
 *       512 for 64-bit storage);
 *   "u_passes" 0: column pairs per workgroup building U, 1, 2, 4 or 8 (0: 1
 *       for fewer than 8 matrices in a step, 4 from 8 on);
 *   "inner_split" 16: at most this many workgroups per 64-row stripe of the
 *       inner GEMM (0: 8 for 64-bit, 4 for 32-bit storage);
 *   "tile64" 1: the 64-bit trailing GEMM's tile, 64 x 32 (1) or 64 x 64 (2);
 *   "fpanel_rows" 16: rows per wave of the float64 panel kernel, 16 or 32.
 * "gemm_group", "inner_split", "tile64" and "fpanel_rows" also take values per
 * size class of a matrix (largest side < 3072, < 12288, larger) under
 * "<key>.s", "<key>.m", "<key>.l" (0: the global value; "fpanel_rows.m" and
 * ".l" default to 32); a launch uses the class of its largest matrix.
 * Unknown keys and out-of-range values return LINAC_ERR_INVALID. */
int linac_tune_set(const char* key, int64_t value);
int linac_tune_get(const char* key, int64_t* value);
/* Architecture and marketing name of the current device ("gfx950: AMD
 * Instinct MI355X"), NUL-terminated in out[0..n); keys tuning files. */
int linac_device_name(char* out, int n);

/* ---- RREF over GF(p) ----
 * Replaces cuda_row_reduce(matrix, field_characteristic=p) for p > 0
 * (linac/pycuda_row_reduce.py:18-151; dtype rule :41-48).  Entries are reduced
 * mod p first.  p must be prime (checked; composite -> LINAC_ERR_INVALID).
 * linac_rref_u32: p < 2^32.  linac_rref_u64: p < 2^63.
 * pivots_out (host, may be NULL) receives min(rows, cols) int64 slots: the
 * pivot column of each non-zero row, in row order; *rank_out the rank.
 * On return rows [0, rank) hold the RREF's non-zero rows and rows
 * [rank, rows) are zero -- exactly the output of cuda_row_reduce. */
int linac_rref_u32(uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p,
                   uint32_t flags, void* stream, int64_t* pivots_out, int64_t* rank_out);
int linac_rref_u64(uint64_t* a, int64_t rows, int64_t cols, int64_t ld, uint64_t p,
                   uint32_t flags, void* stream, int64_t* pivots_out, int64_t* rank_out);

/* The plain method of linac_rref_u32 on a device matrix with NO allocation:
 * every buffer beyond the caller's matrix (state, pivot lists, row flags, the
 * gather buffer of the final permutation) comes from a static device array of
 * 32 MiB that each GPU holds until the process ends.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * Requires linac_init(device) first (which resolves the array's address on the
 * device), LINAC_DEVICE_PTR, LINAC_METHOD_PLAIN or _AUTO (plain runs), no
 * LINAC_PROFILE, and rows * ld * 4 + 17 * min(rows, cols) + rows + ~1 KiB
 * <= 32 MiB (LINAC_ERR_INVALID otherwise).  Results equal linac_rref_u32 with
 * LINAC_METHOD_PLAIN.  Calls share the device's array: calls that overlap on
 * one GPU race (the library serialises its own calls per device). */
int linac_rref_u32_noalloc(uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p,
                           uint32_t flags, void* stream, int64_t* pivots_out, int64_t* rank_out);

/* ---- batches of GF(p) matrices (blocked method, 2 < p < 2^63) ----
 * A batch advances all its unfinished matrices together, one 256-column panel
 * per step, in shared kernel launches (one launch per kernel for the whole
 * batch and storage width: uint32 matrices for p <= 4278124287, uint64
 * matrices for 2^32 <= p < 2^63, both in one batch; ids are shared).  Matrices may be added between steps and taken out as soon as they
 * are finished, so a serving loop can keep the device busy while requests
 * arrive and leave (continuous batching).  Each matrix lives in its own device
 * allocation of linac_batch_footprint_u32(rows, cols) bytes; with a non-zero
 * cap, linac_batch_add_u32 refuses (LINAC_ERR_NOMEM) a matrix that would
 * exceed it.  A batch object must be used from one thread at a time.
 * No counterpart in the reference (batching is new); results equal
 * linac_rref_u32 on each matrix. */
typedef struct linac_batch linac_batch;
int linac_batch_footprint_u32(int64_t rows, int64_t cols, uint64_t* bytes_out);
int linac_batch_create(linac_batch** out, uint64_t mem_cap_bytes);
int linac_batch_destroy(linac_batch* b);
int linac_batch_memory(linac_batch* b, uint64_t* in_use_out, uint64_t* cap_out);
/* Copies the matrix (host, or device with LINAC_DEVICE_PTR) into the batch.
 * With LINAC_DEVICE_PTR | LINAC_IN_PLACE the batch works on the caller's
 * device matrix itself (no copy; the footprint then excludes the matrix,
 * linac_batch_footprint_u32 minus rows*round_up(cols,64)*4 rounded to 256):
 * the buffer must stay allocated until the matrix is taken, and holds its RREF
 * afterwards. */
int linac_batch_add_u32(linac_batch* b, const uint32_t* a, int64_t rows, int64_t cols, int64_t ld, uint32_t p,
                        uint32_t flags, int64_t* id_out);
/* The same for 64-bit storage, 2^32 <= p < 2^63 (footprint: _u64; the matrix
 * part is rows*round_up(cols,64)*8 rounded to 256). */
int linac_batch_footprint_u64(int64_t rows, int64_t cols, uint64_t* bytes_out);
int linac_batch_add_u64(linac_batch* b, const uint64_t* a, int64_t rows, int64_t cols, int64_t ld, uint64_t p,
                        uint32_t flags, int64_t* id_out);
/* One panel for every unfinished matrix; *unfinished_out counts those left. */
int linac_batch_step(linac_batch* b, int64_t* unfinished_out);
int linac_batch_run(linac_batch* b);
/* 1 when finished, 0 while running. */
int linac_batch_status(linac_batch* b, int64_t id);
/* Writes the RREF of a finished matrix to out (host, or device with
 * LINAC_DEVICE_PTR; may be NULL), its pivot columns and rank, and releases it. */
int linac_batch_take_u32(linac_batch* b, int64_t id, uint32_t* out, int64_t ld_out, uint32_t flags,
                         int64_t* pivots_out, int64_t* rank_out);
int linac_batch_take_u64(linac_batch* b, int64_t id, uint64_t* out, int64_t ld_out, uint32_t flags,
                         int64_t* pivots_out, int64_t* rank_out);
/* Many finished matrices at once: as linac_batch_take_u32 for each ids[i]
 * (outs[i] / pivots_out[i] may be NULL, as may the arrays themselves), with two
 * host round trips in all instead of several per matrix. */
int linac_batch_take_many_u32(linac_batch* b, const int64_t* ids, int64_t count, uint32_t* const* outs,
                              const int64_t* ld_outs, uint32_t flags, int64_t* const* pivots_out,
                              int64_t* ranks_out);
int linac_batch_take_many_u64(linac_batch* b, const int64_t* ids, int64_t count, uint64_t* const* outs,
                              const int64_t* ld_outs, uint32_t flags, int64_t* const* pivots_out,
                              int64_t* ranks_out);
/* One-shot: RREF of `count` matrices in place (each with its own shape and
 * prime); pivots_out may be NULL or hold per-matrix pointers (or NULLs). */
int linac_rref_batch_u32(uint32_t* const* mats, const int64_t* rows, const int64_t* cols, const int64_t* lds,
                         const uint32_t* primes, int64_t count, uint32_t flags, int64_t* const* pivots_out,
                         int64_t* ranks_out);

/* ---- RREF over float64 / complex128 ----
 * Replaces cuda_row_reduce(matrix, 0) (linac/pycuda_row_reduce.py:49-56,87-131;
 * kernels linac/row_reduce.cu:201-270,287-342,404-418): rows are divided once by
 * their max |x|, the pivot of a column is the unused row of largest |x|, and it
 * is accepted iff |x| > tol (the reference uses 1e-9).  c128 arrays hold
 * interleaved (re, im) doubles; ld counts complex elements. */
int linac_rref_f64(double* a, int64_t rows, int64_t cols, int64_t ld, double tol, uint32_t flags, void* stream,
                   int64_t* pivots_out, int64_t* rank_out);
int linac_rref_c128(double* a, int64_t rows, int64_t cols, int64_t ld, double tol, uint32_t flags, void* stream,
                    int64_t* pivots_out, int64_t* rank_out);
/* With LINAC_REFINE (blocked method only; AUTO then runs blocked, PLAIN is
 * LINAC_ERR_INVALID) the call also keeps every panel's factors and runs two
 * passes of residual-based refinement on the non-pivot columns: the residual
 * A[:, NP] - A[:, P] X of the original matrix in double-double arithmetic,
 * replayed through the factors, added to X.  The pivot decisions (and so the
 * rank and pivot columns) are those of the unrefined elimination.  No
 * counterpart in the reference (its float path has no refinement,
 * linac/row_reduce.cu:138-196).  Extra device memory: about 2 x the matrix.
 * linac_refine_last reports this thread's last refined call:
 *   out[0] passes run, out[1] max|correction| / max|X| of the first pass (an
 *   estimate of the unrefined result's relative error), out[2] the same for
 *   the last pass, out[3] the number of non-pivot columns, out[4] 1 when a
 *   LINAC_METHOD_MIXED call was redone with FP64 updates (below).
 * Writes min(n, 5) values and returns how many were written (zeros after a
 * call without LINAC_REFINE). */
int linac_refine_last(double* out, int n);
/* LINAC_METHOD_MIXED (linac_rref_f64 only; opt-in): the blocked sweep with its
 * trailing updates on FP32 matrix cores -- operands rounded to FP32, products
 * summed in FP32 and added to the FP64 matrix -- while pivot search, row
 * scaling and the panel factors stay FP64; then refinement passes as with
 * LINAC_REFINE until a pass corrects less than 2^-51 relative (at most 10,
 * and no longer once a pass corrects more than half as much as the one
 * before).  If the correction is still above 2^-51 the call starts again from
 * the original matrix with FP64 updates and two refinement passes (as
 * LINAC_METHOD_BLOCKED | LINAC_REFINE), so the result has FP64 accuracy
 * either way; linac_refine_last reports the passes and whether this
 * happened.  The FP32 updates perturb the matrix by about 2^-24 relative,
 * which the pivot test sees: intended for full-rank systems with cond(A) well
 * below 2^24, where it needs no rerun.  No counterpart in the reference (its
 * float path is FP64 throughout, linac/row_reduce.cu:138-196). */

/* ---- resumable float RREF (blocked method, device matrix) ----
 * The same elimination as linac_rref_f64 / linac_rref_c128 with
 * LINAC_DEVICE_PTR and LINAC_METHOD_BLOCKED, split into calls that each run
 * at most max_panels of the 64-column panels (and wait for them), so that a
 * long solve can be interleaved with other work -- the serving loop runs
 * float requests a few panels at a time between batch steps.  finish runs
 * any remaining panels, writes the RREF in place, the pivot columns and rank,
 * and releases the job; destroy abandons it.  Workspace: a device buffer of
 * linac_float_job_workspace bytes (256-byte aligned) that the caller keeps
 * until finish / destroy; with workspace NULL the device context's is used
 * and only one such job may be open per device.  flags must hold
 * LINAC_DEVICE_PTR.  linac_shutdown releases open jobs (destroy on such a
 * handle is a no-op).  No counterpart in the reference, whose calls solve one
 * matrix at once (linac/pycuda_row_reduce.py:94-131). */
typedef struct linac_float_job linac_float_job;
int linac_float_job_workspace(int64_t rows, int64_t cols, int is_complex, int64_t* bytes_out);
int linac_float_job_create(linac_float_job** out, double* a, int64_t rows, int64_t cols, int64_t ld, int is_complex,
                           double tol, uint32_t flags, void* workspace, int64_t workspace_bytes);
int linac_float_job_step(linac_float_job* job, int64_t max_panels, int64_t* remaining_out);
int linac_float_job_finish(linac_float_job* job, int64_t* pivots_out, int64_t* rank_out);
int linac_float_job_destroy(linac_float_job* job);

/* ---- monomial-evaluation matrix loader ----
 * Replaces cuda_load_matrix(bases, lindices, shape, field_characteristic)
 * (linac/pycuda_matrix_loader.py:28-124; kernel linac/matrix_loader.cu:88-112):
 *   out[r, c] = prod_{t < degree} bases[r, lindices[c, t]]   (mod p for u32/u64)
 * bases: rows x basis_len (row-major), lindices: cols x degree uint32,
 * out: rows x cols (row-major, contiguous).  Bases are reduced mod p first.
 * 1 <= degree <= 32, 1 <= basis_len <= 4096; every index must be < basis_len
 * (checked; LINAC_ERR_INVALID otherwise).  Systems sharing lindices are one
 * call with their bases stacked row-wise.  c128 arrays are interleaved
 * (re, im) doubles. */
int linac_load_matrix_u32(uint32_t* out, const uint32_t* bases, const uint32_t* lindices, int64_t rows,
                          int64_t cols, int64_t basis_len, int64_t degree, uint32_t p, uint32_t flags,
                          void* stream);
int linac_load_matrix_u64(uint64_t* out, const uint64_t* bases, const uint32_t* lindices, int64_t rows,
                          int64_t cols, int64_t basis_len, int64_t degree, uint64_t p, uint32_t flags,
                          void* stream);
int linac_load_matrix_f64(double* out, const double* bases, const uint32_t* lindices, int64_t rows, int64_t cols,
                          int64_t basis_len, int64_t degree, uint32_t flags, void* stream);
int linac_load_matrix_c128(double* out, const double* bases, const uint32_t* lindices, int64_t rows, int64_t cols,
                           int64_t basis_len, int64_t degree, uint32_t flags, void* stream);

/* ------------------------------------------------------------------------
 * Exact modular GEMM with index lists, on the int8-limb matrix cores:
 *   C[i, j] <- (C[i, j] - sum_{t<k} A[i, a_cols[t]] * B[b_rows[t], j]) mod p
 * for i < m, j < n.  C (row stride ldc), A (m x a_ncols, stride lda) and B
 * (b_nrows x n, stride ldb) are device matrices with canonical entries in
 * [0, p) (flags must include LINAC_DEVICE_PTR); a_cols / b_rows are host
 * arrays of k indices, or NULL for 0..k-1.  C may be the same matrix as A or
 * B: all operands are captured before C is written.  2 < p <= 4278124287.
 * Synchronous.  No counterpart in the reference, which re-reduces a whole
 * stacked matrix instead (linac/vector_spaces/over_function_fields.py:74-101);
 * this is the update step of the resident incremental RREF
 * (linac_amd/incremental.py).
 * ------------------------------------------------------------------------ */
int linac_gemm_sub_u32(uint32_t* C, int64_t ldc, const uint32_t* A, int64_t lda, int64_t a_ncols, const int64_t* a_cols,
                       const uint32_t* B, int64_t ldb, int64_t b_nrows, const int64_t* b_rows, int64_t m, int64_t n,
                       int64_t k, uint32_t p, uint32_t flags);

/* ------------------------------------------------------------------------
 * Sparse systems from disk.  Replaces the dense host construction of
 * linac/sparse_matrix_tools.py:37-44 (matrix_from_coo) for the GPU path:
 * out (a device matrix, rows x cols, row stride ld elements) is zeroed and
 * out[row_idx[i], col_idx[i]] = values[i] for i < nnz.  Values must already
 * be reduced modulo the field's prime.  The index and value arrays are host
 * memory, or device memory with LINAC_DEVICE_PTR.  An index outside the
 * matrix is LINAC_ERR_INVALID (checked on the host for host arrays, by the
 * kernel otherwise).  Synchronous.
 * ------------------------------------------------------------------------ */
int linac_coo_scatter_u32(uint32_t* out, int64_t ld, int64_t rows, int64_t cols, const int64_t* row_idx,
                          const int64_t* col_idx, const uint32_t* values, int64_t nnz, uint32_t flags);
int linac_coo_scatter_u64(uint64_t* out, int64_t ld, int64_t rows, int64_t cols, const int64_t* row_idx,
                          const int64_t* col_idx, const uint64_t* values, int64_t nnz, uint32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* LINAC_AMD_H */
