// Exact modular GEMM on the int8 matrix cores of gfx950.
//
//   C[m, n] <- (C[m, n] - sum_k L[m, k] * B[k, n]) mod p        (p odd, p <= 4278124287)
//
// Every operand entry x in [0, p) is first mapped to its centred representative
// c in (-p/2, p/2] and written as four balanced base-256 digits
//   c = d0 + 256 d1 + 65536 d2 + 2^24 d3,   d_i in [-128, 127],
// which is possible for |c| <= 127 * (1 + 2^8 + 2^16 + 2^24) = 2139062143,
// i.e. for every p <= 4278124287.  The operands are stored as four int8 "limb
// planes" (plane l holds digit l of every entry, K contiguous).
//
// With D_{l1,l2} = sum_k L_l1[m, k] * B_l2[k, n] (one MFMA chain per limb pair),
//   sum_k L[m,k] B[k,n] = sum_{s=0..6} 256^s S_s,   S_s = sum_{l1+l2=s} D_{l1,l2}.
// Exactness of the int32 accumulators: |digit products| <= 2^14, so
// |S_s| <= 4 * K * 2^14 <= 2^25 for K <= 512 (2^31 would allow K < 2^15).
//
// The seven S_s are folded modulo p in double precision, where every value is
// an integer below 2^53 and therefore exact:
//   P0 = S0 + 256 S1, P1 = S2 + 256 S3, P2 = S4 + 256 S5, P3 = S6  (|P| < 2^34)
//   r = P3;  r = red(r * 2^16 + P2);  r = red(r * 2^16 + P1);  r = r * 2^16 + P0
// with red(x) = x - rint(x / p) * p (an FMA, exact because the result is a
// small integer); |r| stays below ~1.5 p < 2^32, so r * 2^16 + P < 2^49.
// One floor-based reduction of C - r (|C - r| < 2^50) brings it into [0, p):
// q = floor((C - r) / p) is off by at most one, which the two corrections fix.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace linac {

typedef int v4i __attribute__((ext_vector_type(4)));

// Largest prime handled by the limb GEMM (centred values fit 4 balanced digits).
constexpr uint64_t LIMB_P_MAX = 4278124287ull;

// Loads through the global address space.  Pointers read from a descriptor
// are generic, and generic (flat) loads count against lgkmcnt as well as
// vmcnt, so the wait for an LDS fragment would also wait for every global
// load in flight (the next K stage's prefetch); global loads count in vmcnt
// only.
template <class V>
__device__ __forceinline__ V gld(const void* p) {
    return *(const __attribute__((address_space(1))) V*)p;
}
// Stores through the global address space (a generic store would make the
// next wait for a scalar load wait for it as well).
template <class V>
__device__ __forceinline__ void gst(void* p, V v) {
    *(__attribute__((address_space(1))) V*)p = v;
}

__host__ __device__ __forceinline__ int32_t centred(uint32_t v, uint32_t p) {
    return v > (p - 1) / 2 ? (int32_t)((int64_t)v - (int64_t)p) : (int32_t)v;
}

// Balanced base-256 digits of c (|c| <= 2139062143); d[3] is the top digit.
__host__ __device__ __forceinline__ void balanced_digits(int32_t c, int8_t d[4]) {
    int32_t x = c;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
        const int8_t dl = (int8_t)(x & 0xFF);
        d[l] = dl;
        x = (x - dl) >> 8;
    }
    d[3] = (int8_t)x;
}

__host__ __device__ __forceinline__ double red_rint(double x, double p, double pinv) {
    const double q = rint(x * pinv);
    return fma(-q, p, x);
}

// (cval - sum_s 256^s S_s) mod p, or (- sum) mod p when cval is not used.
__host__ __device__ __forceinline__ uint32_t fold_sub(uint32_t cval, const int32_t s[7], double p, double pinv) {
    const double P0 = (double)s[0] + 256.0 * (double)s[1];
    const double P1 = (double)s[2] + 256.0 * (double)s[3];
    const double P2 = (double)s[4] + 256.0 * (double)s[5];
    double r = (double)s[6];
    r = red_rint(fma(r, 65536.0, P2), p, pinv);
    r = red_rint(fma(r, 65536.0, P1), p, pinv);
    r = fma(r, 65536.0, P0);  // |r| < 2^48: reduced together with C below
    // |t| < 2^48 + p, so t / p < 2^18 rounds with an error far below 1/2 and
    // the nearest-quotient remainder lies within [-p/2, p/2] up to ties, well
    // inside (-p, p): one correction brings it to [0, p)
    const double t = red_rint((double)cval - r, p, pinv);
    return (uint32_t)(t < 0.0 ? t + p : t);
}

// fold_sub for sums of at most 32 limb products per diagonal entry (the
// inner GEMM: k <= 32 pivots), where |S_s| <= 4 * 32 * 2^14 = 2^21, so the
// pairs S_2i + 256 S_2i+1 are exact in int32 (|.| < 2^30) and only four values
// go through double precision.
__host__ __device__ __forceinline__ uint32_t fold_sub_k32(uint32_t cval, const int32_t s[7], double p, double pinv) {
    const int32_t Q0 = s[0] + s[1] * 256, Q1 = s[2] + s[3] * 256, Q2 = s[4] + s[5] * 256;
    double r = (double)s[6];
    r = red_rint(fma(r, 65536.0, (double)Q2), p, pinv);
    r = red_rint(fma(r, 65536.0, (double)Q1), p, pinv);
    r = fma(r, 65536.0, (double)Q0);
    const double t = red_rint((double)cval - r, p, pinv);  // within (-p, p), as in fold_sub
    return (uint32_t)(t < 0.0 ? t + p : t);
}

// ---------------------------------------------------------------------------
// 64-bit storage: 2^32 < p < 2^63.  Centred values (|c| < 2^62) take eight
// balanced base-256 digits (the top one within [-65, 65]), so a modular
// product needs 64 limb products whose sums fall on 15 diagonals
//   S_s = sum_{l1+l2=s} D_{l1,l2},  |S_s| <= 8 * K * 2^14 = 2^26  (K <= 512).
// The fold groups them in base-2^24 triples
//   T_i = S_3i + 2^8 S_3i+1 + 2^16 S_3i+2   (|T_i| < 2^43, exact in int64 and double)
// and runs Horner r <- r 2^24 + T_i with a reduction at every step: the
// quotient q = rint((r 2^24 + T_i) / p) comes from double precision (|q| <
// 2^24 once |r| <= p/2, since |r| stays within (1/2 + 2^-27) p), and the
// remainder is formed exactly in wrapping 64-bit integers, which is correct
// because its true value is below p in magnitude.
// ---------------------------------------------------------------------------

constexpr uint64_t LIMB64_P_MIN = 1ull << 32;   // smaller primes use 32-bit storage
constexpr uint64_t LIMB64_P_MAX = 1ull << 63;   // exclusive; the field arithmetic needs p < 2^63

__host__ __device__ __forceinline__ int64_t centred(uint64_t v, uint64_t p) {
    return v > (p - 1) / 2 ? (int64_t)(v - p) : (int64_t)v;
}

__host__ __device__ __forceinline__ void balanced_digits(int64_t c, int8_t d[8]) {
    int64_t x = c;
#pragma unroll
    for (int l = 0; l < 7; ++l) {
        const int8_t dl = (int8_t)(x & 0xFF);
        d[l] = dl;
        x = (x - dl) >> 8;
    }
    d[7] = (int8_t)x;
}

// (cval - sum_{s<15} 256^s S_s) mod p for 2^32 <= p < 2^63.
__host__ __device__ __forceinline__ uint64_t fold_sub64(uint64_t cval, const int32_t s[15], uint64_t p, double pinv) {
    int64_t T[5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
        T[i] = (int64_t)s[3 * i] + (int64_t)s[3 * i + 1] * 256 + (int64_t)s[3 * i + 2] * 65536;
    int64_t r = T[4];
    r -= (int64_t)(int32_t)rint((double)r * pinv) * (int64_t)p;  // |T_4| / p < 2^11
#pragma unroll
    for (int i = 3; i >= 0; --i) {
        const double xd = fma((double)r, 16777216.0, (double)T[i]);
        const int32_t q = (int32_t)rint(xd * pinv);
        r = (int64_t)(((uint64_t)r << 24) + (uint64_t)T[i] - (uint64_t)((int64_t)q * (int64_t)p));
    }
    const uint64_t ur = r < 0 ? (uint64_t)(r + (int64_t)p) : (uint64_t)r;
    return cval >= ur ? cval - ur : cval + (p - ur);
}

}  // namespace linac
