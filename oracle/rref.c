/*
 * C restatement of linac's CPU row reduction (TEST INFRASTRUCTURE ONLY).
 *
 * It follows linac/row_reduce.py:16-78 with the arguments the reference's own
 * callers pass for finite fields (pivoting=1, scaling=False, threshold=0,
 * prime=p; linac/vector_spaces/over_number_fields.py:109) and, for floats, with
 * the defaults that match the GPU float path (pivoting=1, scaling=True,
 * threshold=1e-9; linac/row_reduce.cu:55,201-270,404-418):
 *
 *   while i < rows and j < cols:                       row_reduce.py:40
 *     pivot_row = first argmax |A[i:, j]|              row_reduce.py:45-47
 *     swap rows i, pivot_row                           row_reduce.py:47
 *     if |A[i, j]| > threshold:                        row_reduce.py:60
 *       A[i, :] *= inverse(A[i, j])                    row_reduce.py:61-67
 *       A[r, j:] -= A[r, j] * A[i, j:]  for r != i     row_reduce.py:68-70
 *       i += 1; j += 1
 *     else: j += 1
 *
 * Finite-field values are kept canonical in [0, p) after every step, which is
 * what the per-step "% prime" of row_reduce.py:75-76 does.  Storage is uint64
 * with unsigned __int128 products, so any p < 2^63 works (the int64 numpy
 * reference is limited to p < 2^31.5).  The elimination over rows is split
 * over OpenMP threads when compiled with -fopenmp; the result does not depend
 * on the thread count.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef unsigned __int128 u128;

/* a*b mod p.  For p < 2^32 the product fits 64 bits and is reduced by Barrett
 * with mu = floor((2^64-1)/p): q = hi64(x*mu) is at most 2 below floor(x/p).
 * For larger p the 128-bit product is reduced with the compiler's % . */
static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p, uint64_t mu) {
    if (p < (1ULL << 32)) {
        uint64_t x = a * b;
        uint64_t q = (uint64_t)(((u128)x * mu) >> 64);
        uint64_t r = x - q * p;
        while (r >= p) r -= p;
        return r;
    }
    return (uint64_t)(((u128)a * b) % p);
}

/* inverse by the extended Euclidean algorithm; returns 0 when gcd != 1 */
static uint64_t invmod(uint64_t a, uint64_t p) {
    __int128 r0 = p, r1 = a % p, s0 = 0, s1 = 1;
    while (r1 != 0) {
        __int128 q = r0 / r1, t;
        t = r0 - q * r1; r0 = r1; r1 = t;
        t = s0 - q * s1; s0 = s1; s1 = t;
    }
    if (r0 != 1) return 0;
    s0 %= (__int128)p;
    if (s0 < 0) s0 += p;
    return (uint64_t)s0;
}

/* One definition per storage type: uint64 (any p < 2^63) and uint32
 * (p < 2^32, half the memory traffic). */
#define DEFINE_RREF_MODP(T, SUFFIX)                                                   \
static void swap_rows_##SUFFIX(T* a, int64_t cols, int64_t r0, int64_t r1) {          \
    if (r0 == r1) return;                                                             \
    for (int64_t c = 0; c < cols; ++c) {                                              \
        T t = a[r0 * cols + c]; a[r0 * cols + c] = a[r1 * cols + c]; a[r1 * cols + c] = t; \
    }                                                                                 \
}                                                                                     \
int oracle_rref_modp_##SUFFIX(T* a, int64_t rows, int64_t cols, uint64_t p,          \
                              int64_t* pivcols, int64_t* rank) {                      \
    const uint64_t mu = UINT64_MAX / p;                                               \
    for (int64_t x = 0; x < rows * cols; ++x) a[x] = (T)(a[x] % p);                   \
    int64_t i = 0, j = 0;                                                             \
    while (i < rows && j < cols) {                                                    \
        int64_t prow = i;                                                             \
        uint64_t best = a[i * cols + j];                                              \
        for (int64_t r = i + 1; r < rows; ++r) {                                      \
            uint64_t v = a[r * cols + j];                                             \
            if (v > best) { best = v; prow = r; }                                     \
        }                                                                             \
        swap_rows_##SUFFIX(a, cols, i, prow);                                         \
        uint64_t head = a[i * cols + j];                                              \
        if (head != 0) {                                                              \
            uint64_t inv = invmod(head, p);                                           \
            if (inv == 0) return -1;                                                  \
            T* pr = a + i * cols;                                                     \
            for (int64_t c = j; c < cols; ++c) pr[c] = (T)mulmod(pr[c], inv, p, mu);  \
            _Pragma("omp parallel for schedule(static)")                              \
            for (int64_t r = 0; r < rows; ++r) {                                      \
                if (r == i) continue;                                                 \
                T* row = a + r * cols;                                                \
                uint64_t f = row[j];                                                  \
                if (f == 0) continue;                                                 \
                for (int64_t c = j; c < cols; ++c) {                                  \
                    uint64_t m = mulmod(f, pr[c], p, mu);                             \
                    uint64_t v = row[c];                                              \
                    row[c] = (T)(v >= m ? v - m : v + (p - m));                       \
                }                                                                     \
            }                                                                         \
            if (pivcols) pivcols[i] = j;                                              \
            i += 1; j += 1;                                                           \
        } else {                                                                      \
            j += 1;                                                                   \
        }                                                                             \
    }                                                                                 \
    if (rank) *rank = i;                                                              \
    return 0;                                                                         \
}

/* Returns 0 on success, -1 when a head is not invertible (composite p), the
 * analogue of the ZeroDivisionError at row_reduce.py:63-64.  pivcols (may be
 * NULL) receives the pivot column of each non-zero row; *rank the count. */
DEFINE_RREF_MODP(uint64_t, u64)
DEFINE_RREF_MODP(uint32_t, u32)

/* Float path: one-shot row scaling, partial pivoting on |x|, threshold. */
int oracle_rref_f64(double* a, int64_t rows, int64_t cols, double threshold,
                    int64_t* pivcols, int64_t* rank) {
    for (int64_t r = 0; r < rows; ++r) {
        double s = 0.0;
        for (int64_t c = 0; c < cols; ++c) s = fmax(s, fabs(a[r * cols + c]));
        for (int64_t c = 0; c < cols; ++c) a[r * cols + c] /= s;
    }
    int64_t i = 0, j = 0;
    while (i < rows && j < cols) {
        int64_t prow = i;
        double best = fabs(a[i * cols + j]);
        for (int64_t r = i + 1; r < rows; ++r) {
            double v = fabs(a[r * cols + j]);
            if (v > best) { best = v; prow = r; }
        }
        if (prow != i) {
            for (int64_t c = 0; c < cols; ++c) {
                double t = a[i * cols + c]; a[i * cols + c] = a[prow * cols + c]; a[prow * cols + c] = t;
            }
        }
        double head = a[i * cols + j];
        if (fabs(head) > threshold) {
            double* pr = a + i * cols;
            for (int64_t c = 0; c < cols; ++c) pr[c] = pr[c] / head;
            #pragma omp parallel for schedule(static)
            for (int64_t r = 0; r < rows; ++r) {
                if (r == i) continue;
                double* row = a + r * cols;
                double f = row[j];
                for (int64_t c = j; c < cols; ++c) row[c] = row[c] - f * pr[c];
            }
            if (pivcols) pivcols[i] = j;
            i += 1; j += 1;
        } else {
            j += 1;
        }
    }
    if (rank) *rank = i;
    return 0;
}

/* Complex path (interleaved re, im), same steps with |z| for pivoting. */
int oracle_rref_c128(double* a_ri, int64_t rows, int64_t cols, double threshold,
                     int64_t* pivcols, int64_t* rank) {
    double complex* a = (double complex*)a_ri;
    for (int64_t r = 0; r < rows; ++r) {
        double s = 0.0;
        for (int64_t c = 0; c < cols; ++c) s = fmax(s, cabs(a[r * cols + c]));
        for (int64_t c = 0; c < cols; ++c) a[r * cols + c] /= s;
    }
    int64_t i = 0, j = 0;
    while (i < rows && j < cols) {
        int64_t prow = i;
        double best = cabs(a[i * cols + j]);
        for (int64_t r = i + 1; r < rows; ++r) {
            double v = cabs(a[r * cols + j]);
            if (v > best) { best = v; prow = r; }
        }
        if (prow != i) {
            for (int64_t c = 0; c < cols; ++c) {
                double complex t = a[i * cols + c]; a[i * cols + c] = a[prow * cols + c]; a[prow * cols + c] = t;
            }
        }
        double complex head = a[i * cols + j];
        if (cabs(head) > threshold) {
            double complex* pr = a + i * cols;
            for (int64_t c = 0; c < cols; ++c) pr[c] = pr[c] / head;
            #pragma omp parallel for schedule(static)
            for (int64_t r = 0; r < rows; ++r) {
                if (r == i) continue;
                double complex* row = a + r * cols;
                double complex f = row[j];
                for (int64_t c = j; c < cols; ++c) row[c] = row[c] - f * pr[c];
            }
            if (pivcols) pivcols[i] = j;
            i += 1; j += 1;
        } else {
            j += 1;
        }
    }
    if (rank) *rank = i;
    return 0;
}

/* Monomial evaluation of linac/matrix_loader.cu:88-112 / the Python-int
 * fallback of linac/pycuda_matrix_loader.py:221-223:
 *   out[r, c] = prod_t bases[r, lindices[c, t]]  (mod p). */
void oracle_load_matrix_modp_u64(uint64_t* out, const uint64_t* bases, const uint32_t* lindices,
                                 int64_t rows, int64_t cols, int64_t basis_len, int64_t degree,
                                 uint64_t p) {
    const uint64_t mu = UINT64_MAX / p;
    #pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < rows; ++r) {
        for (int64_t c = 0; c < cols; ++c) {
            uint64_t v = bases[r * basis_len + lindices[c * degree]] % p;
            for (int64_t t = 1; t < degree; ++t)
                v = mulmod(v, bases[r * basis_len + lindices[c * degree + t]] % p, p, mu);
            out[r * cols + c] = v;
        }
    }
}
