// Exact prime-field arithmetic for gfx950 kernels.
//
// Elements are stored canonical in [0, p).  Products go through Montgomery
// reduction with R = 2^32 (FieldU32) or R = 2^64 (FieldU64) so that the prime
// is a runtime argument (one compiled kernel serves every prime) and no
// integer division runs on the device.
//
// Montgomery REDC of t = a*b < p*2^w (w = 32 or 64):
//   m = (t mod 2^w) * (-p^{-1}) mod 2^w,  u = (t + m*p) / 2^w  <  2p.
// Since t + m*p == 0 (mod 2^w), the carry out of the low word is exactly
// (t_lo != 0), so u = t_hi + hi(m*p) + (t_lo != 0) without forming t + m*p.
// u < 2p fits in w+1 bits; FieldU32 keeps it in 64 bits, FieldU64 needs
// p < 2^63 so that 2p < 2^64.
//
// A multiplier f is pre-converted to Montgomery form (f*R mod p) once per
// row, after which mont(fR, x) = f*x mod p with a single REDC per element.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace linac {

// Modular inverse of a (0 < a < p, p odd, gcd = 1) by the binary extended
// Euclidean algorithm: only shifts, adds and compares, so the chain stays
// short on a single lane.  Invariants: x1*a == u, x2*a == v (mod p).
template <class U, class W>
__host__ __device__ inline U binary_inverse(U a, U p) {
    U u = a, v = p, x1 = 1, x2 = 0;
    while (u != 1 && v != 1) {
        while ((u & 1) == 0) {
            u >>= 1;
            x1 = (x1 & 1) ? (U)(((W)x1 + p) >> 1) : (U)(x1 >> 1);
        }
        while ((v & 1) == 0) {
            v >>= 1;
            x2 = (x2 & 1) ? (U)(((W)x2 + p) >> 1) : (U)(x2 >> 1);
        }
        if (u >= v) {
            u -= v;
            x1 = x1 >= x2 ? x1 - x2 : (U)(x1 + (p - x2));
        } else {
            v -= u;
            x2 = x2 >= x1 ? x2 - x1 : (U)(x2 + (p - x1));
        }
    }
    return u == 1 ? x1 : x2;
}

struct FieldU32 {
    using T = uint32_t;
    uint32_t p;
    uint32_t pneg;  // -p^{-1} mod 2^32
    uint32_t r2;    // 2^64 mod p

    __device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b) const {
        const uint32_t t_lo = a * b;
        const uint32_t t_hi = __umulhi(a, b);
        const uint32_t m = t_lo * pneg;
        const uint64_t u = (uint64_t)t_hi + __umulhi(m, p) + (t_lo != 0u);
        return (uint32_t)(u >= p ? u - p : u);
    }
    __device__ __forceinline__ uint32_t sub(uint32_t a, uint32_t b) const {
        return a >= b ? a - b : a + (p - b);
    }
    __device__ __forceinline__ uint32_t to_mont(uint32_t a) const { return mont(a, r2); }
    // Normal-form multiplier X with mont(a, X) = a * h^{-1} * R (the
    // Montgomery form of the elimination factor a / h).
    __device__ __forceinline__ uint32_t factor_of_head(uint32_t h) const {
        const uint32_t inv = binary_inverse<uint32_t, uint64_t>(h, p);
        return to_mont(to_mont(inv));
    }
    __device__ __forceinline__ uint32_t inverse(uint32_t h) const {
        return binary_inverse<uint32_t, uint64_t>(h, p);
    }
    // a * b mod p for two normal-form operands.
    __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) const { return mont(to_mont(a), b); }
    __device__ __forceinline__ uint32_t reduce(uint32_t a) const { return a >= p ? a % p : a; }
};

struct FieldU64 {
    using T = uint64_t;
    uint64_t p;
    uint64_t pneg;  // -p^{-1} mod 2^64
    uint64_t r2;    // 2^128 mod p

    __device__ __forceinline__ uint64_t mont(uint64_t a, uint64_t b) const {
        const uint64_t t_lo = a * b;
        const uint64_t t_hi = __umul64hi(a, b);
        const uint64_t m = t_lo * pneg;
        const uint64_t u = t_hi + __umul64hi(m, p) + (t_lo != 0ull);
        return u >= p ? u - p : u;
    }
    __device__ __forceinline__ uint64_t sub(uint64_t a, uint64_t b) const {
        return a >= b ? a - b : a + (p - b);
    }
    __device__ __forceinline__ uint64_t to_mont(uint64_t a) const { return mont(a, r2); }
    __device__ __forceinline__ uint64_t factor_of_head(uint64_t h) const {
        const uint64_t inv = binary_inverse<uint64_t, uint64_t>(h, p);  // p < 2^63: x + p cannot wrap
        return to_mont(to_mont(inv));
    }
    __device__ __forceinline__ uint64_t inverse(uint64_t h) const {
        return binary_inverse<uint64_t, uint64_t>(h, p);  // p < 2^63: x + p cannot wrap
    }
    __device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b) const { return mont(to_mont(a), b); }
    __device__ __forceinline__ uint64_t reduce(uint64_t a) const { return a >= p ? a % p : a; }
};

// GF(2): the only even prime; Montgomery needs an odd modulus.
struct FieldGF2 {
    using T = uint32_t;
    uint32_t p;  // == 2
    __device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b) const { return a & b; }
    __device__ __forceinline__ uint32_t sub(uint32_t a, uint32_t b) const { return a ^ b; }
    __device__ __forceinline__ uint32_t factor_of_head(uint32_t) const { return 1u; }
    __device__ __forceinline__ uint32_t inverse(uint32_t) const { return 1u; }
    __device__ __forceinline__ uint32_t to_mont(uint32_t a) const { return a; }
    __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) const { return a & b; }
    __device__ __forceinline__ uint32_t reduce(uint32_t a) const { return a & 1u; }
};

// Host-side construction of the Montgomery constants.
inline FieldU32 make_field_u32(uint32_t p) {
    FieldU32 f;
    f.p = p;
    uint32_t inv = 1;  // Newton iteration for p^{-1} mod 2^32 (p odd)
    for (int i = 0; i < 5; ++i) inv *= 2u - p * inv;
    f.pneg = 0u - inv;
    const unsigned __int128 r2 = (((unsigned __int128)1) << 64) % p;
    f.r2 = (uint32_t)r2;
    return f;
}

inline FieldU64 make_field_u64(uint64_t p) {
    FieldU64 f;
    f.p = p;
    uint64_t inv = 1;
    for (int i = 0; i < 6; ++i) inv *= 2ull - p * inv;
    f.pneg = 0ull - inv;
    // 2^128 mod p = ((2^64 mod p)^2) mod p
    const unsigned __int128 r = ((unsigned __int128)1 << 64) % p;
    f.r2 = (uint64_t)((r * r) % p);
    return f;
}

}  // namespace linac
