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

// High words of products, and leading zeros: the device instructions in
// kernels, plain C++ in host code (the CPU exactness tests call the field and
// fold routines on the host).
__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}
__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}
__host__ __device__ __forceinline__ int clz32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clz(x);
#else
    return __builtin_clz(x);
#endif
}
__host__ __device__ __forceinline__ int clz64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clzll(x);
#else
    return __builtin_clzll(x);
#endif
}

// a^{-1} mod p (p an odd prime, 0 < a < p) by the binary extended Euclidean
// algorithm (Knuth, TAOCP vol. 2, 4.5.2, binary gcd with cofactors), under the
// invariants x1 a = u and x2 a = v (mod p): halve the even one of u, v (and its
// cofactor, (x + p) / 2 for odd x), subtract the smaller from the larger, until
// one reaches 1.  At most about 2 log2(p) halvings, each with one subtraction
// at most.  The control flow depends on the data, so it is meant for a value
// held alike in every lane (the loop then runs on the scalar unit), where it
// is several times shorter than the Fermat chain of `inverse`.  a = 0 gives 0.
template <class T>
__host__ __device__ inline T binary_inverse(T a, T p) {
    if (a == 0) return 0;
    T u = a, v = p, x1 = 1, x2 = 0;
    const T half = (p >> 1) + 1;  // (x + p) / 2 = (x >> 1) + half for odd x
    while (u != 1 && v != 1) {
        while ((u & 1) == 0) {
            u >>= 1;
            x1 = (x1 & 1) ? (x1 >> 1) + half : x1 >> 1;
        }
        while ((v & 1) == 0) {
            v >>= 1;
            x2 = (x2 & 1) ? (x2 >> 1) + half : x2 >> 1;
        }
        if (u >= v) {
            u -= v;
            x1 = x1 >= x2 ? x1 - x2 : x1 + (p - x2);
        } else {
            v -= u;
            x2 = x2 >= x1 ? x2 - x1 : x2 + (p - x1);
        }
    }
    return u == 1 ? x1 : x2;
}

// a^{-1} mod p for p < 2^32 (0 < a < p; a = 0 gives 0) by the extended
// Euclidean algorithm in double precision: every remainder and cofactor is an
// integer below 2^33 in magnitude, so the products and differences formed by
// fma are exact.  The quotient comes from a reciprocal refined by one Newton
// step (error far below 1) and is corrected by one in either direction.
// About 0.84 ln(p) ~ 18 steps: a short chain for a value held alike in every
// lane, against about 2 log2(p) data-dependent steps of the binary method.
__host__ __device__ inline uint32_t euclid_inverse_u32(uint32_t a, uint32_t p) {
    double r0 = (double)p, r1 = (double)a, t0 = 0.0, t1 = 1.0;
    while (r1 != 0.0) {
#if defined(__HIP_DEVICE_COMPILE__)
        double inv = __builtin_amdgcn_rcp(r1);
        inv = inv * fma(-r1, inv, 2.0);
#else
        const double inv = 1.0 / r1;
#endif
        double q = floor(r0 * inv);
        double rem = fma(-q, r1, r0);
        if (rem < 0.0) {
            q -= 1.0;
            rem += r1;
        } else if (rem >= r1) {
            q += 1.0;
            rem -= r1;
        }
        const double t = fma(-q, t1, t0);
        r0 = r1;
        r1 = rem;
        t0 = t1;
        t1 = t;
    }
    return (uint32_t)(t0 < 0.0 ? t0 + (double)p : t0);
}

struct FieldU32 {
    using T = uint32_t;
    uint32_t p;
    uint32_t pneg;  // -p^{-1} mod 2^32
    uint32_t r2;    // 2^64 mod p

    __host__ __device__ __forceinline__ uint32_t mont(uint32_t a, uint32_t b) const {
        const uint32_t t_lo = a * b;
        const uint32_t t_hi = mulhi32(a, b);
        const uint32_t m = t_lo * pneg;
        const uint64_t u = (uint64_t)t_hi + mulhi32(m, p) + (t_lo != 0u);
        return (uint32_t)(u >= p ? u - p : u);
    }
    __host__ __device__ __forceinline__ uint32_t sub(uint32_t a, uint32_t b) const {
        return a >= b ? a - b : a + (p - b);
    }
    // a + b mod p without forming a + b, which wraps 32 bits for p > 2^31
    __host__ __device__ __forceinline__ uint32_t add(uint32_t a, uint32_t b) const {
        return a >= p - b ? a - (p - b) : a + b;
    }
    __host__ __device__ __forceinline__ uint32_t to_mont(uint32_t a) const { return mont(a, r2); }
    // h^{-1} = h^(p-2) (Fermat), square-and-multiply in the Montgomery domain:
    // a fixed, branch-uniform sequence of about 2 log2(p) products.
    __host__ __device__ __forceinline__ uint32_t inverse(uint32_t h) const {
        const uint32_t e = p - 2u;
        const uint32_t hm = to_mont(h);
        uint32_t r = to_mont(1u);
        for (int bit = 31 - clz32(e); bit >= 0; --bit) {
            r = mont(r, r);
            if ((e >> bit) & 1u) r = mont(r, hm);
        }
        return mont(r, 1u);
    }
    // Normal-form multiplier X with mont(a, X) = a * h^{-1} * R (the
    // Montgomery form of the elimination factor a / h).
    __host__ __device__ __forceinline__ uint32_t factor_of_head(uint32_t h) const {
        return to_mont(to_mont(inverse(h)));
    }
    // (h x - a y) R^{-1} mod p for canonical h, x, a, y: a fraction-free row
    // update with a constant extra factor R^{-1}.  For p < 2^31 the two
    // products share one reduction: h x + (p - a) y < 2 p^2 < p 2^32.
    __host__ __device__ __forceinline__ uint32_t ff_update(uint32_t h, uint32_t x, uint32_t a, uint32_t y) const {
        if (p < 0x80000000u) return ff_update_small(h, x, a, y);
        return sub(mont(h, x), mont(a, y));
    }
    __host__ __device__ __forceinline__ uint32_t ff_update_small(uint32_t h, uint32_t x, uint32_t a,
                                                                 uint32_t y) const {
        const uint64_t t = (uint64_t)h * x + (uint64_t)(p - a) * y;
        const uint32_t t_lo = (uint32_t)t, t_hi = (uint32_t)(t >> 32);
        const uint32_t m = t_lo * pneg;
        const uint32_t u = t_hi + mulhi32(m, p) + (t_lo != 0u);
        return u >= p ? u - p : u;
    }
    // N updates with one pivot row value y and head h, the choice of reduction
    // made once for all of them (so that their independent chains interleave)
    template <int N>
    __host__ __device__ __forceinline__ void ff_update_n(uint32_t h, const uint32_t* x, const uint32_t* a, uint32_t y,
                                                         uint32_t* out) const {
        if (p < 0x80000000u) {
#pragma unroll
            for (int i = 0; i < N; ++i) out[i] = ff_update_small(h, x[i], a[i], y);
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) out[i] = sub(mont(h, x[i]), mont(a[i], y));
        }
    }
    // a * b mod p for two normal-form operands.
    __host__ __device__ __forceinline__ uint32_t mul(uint32_t a, uint32_t b) const { return mont(to_mont(a), b); }
    __host__ __device__ __forceinline__ uint32_t reduce(uint32_t a) const { return a >= p ? a % p : a; }
};

// p = 2^61 - 1: with 2^61 == 1 (mod p), R = 2^64 == 8 and R^{-1} == 2^58, so
// a Montgomery product t R^{-1} is the residue of t (folded at bit 61) rotated
// left by 58 within 61 bits -- no multiplications beyond the product itself.
// m61_redc takes t = hi 2^64 + lo with hi < 2^59 (t < 2^123).
constexpr uint64_t M61 = (1ull << 61) - 1;
__host__ __device__ __forceinline__ uint64_t m61_redc(uint64_t lo, uint64_t hi) {
    uint64_t x = (lo & M61) + (lo >> 61) + (hi << 3);  // < 2^61 + 8 + 2^62
    x = (x & M61) + (x >> 61);                         // < 2^61 + 4 < 2 p
    x = x >= M61 ? x - M61 : x;                        // t mod p
    return ((x & 7ull) << 58) | (x >> 3);              // x 2^58 mod p (x < p: never p itself)
}

// a b mod 2^61 - 1 for a, b < 2^61 (the fold of m61_redc without its rotation)
__host__ __device__ __forceinline__ uint64_t m61_mul(uint64_t a, uint64_t b) {
    const uint64_t lo = a * b, hi = mulhi64(a, b);  // a b < 2^122: hi < 2^58
    uint64_t x = (lo & M61) + (lo >> 61) + (hi << 3);
    x = (x & M61) + (x >> 61);
    return x >= M61 ? x - M61 : x;
}
// a^(p-2) mod p = a^{-1} for p = 2^61 - 1 (a = 0 gives 0): with x_k = a^(2^k - 1),
// x_{j+k} = x_j^(2^k) x_k, build x_59 from x_1, x_2, x_3, x_6, x_12, x_24, x_48,
// x_56 (with x_8 = x_6^4 x_2), x_58, x_59, then a^(2^61 - 3) = x_59^4 a: 62
// squarings and 11 products, a fixed sequence for a value held alike in every
// lane (a third of the steps of the binary extended GCD, none of its branches).
__host__ __device__ inline uint64_t m61_inverse(uint64_t a) {
    auto sq = [](uint64_t x, int k) {
        for (int i = 0; i < k; ++i) x = m61_mul(x, x);
        return x;
    };
    const uint64_t x1 = a;
    const uint64_t x2 = m61_mul(sq(x1, 1), x1);
    const uint64_t x3 = m61_mul(sq(x2, 1), x1);
    const uint64_t x6 = m61_mul(sq(x3, 3), x3);
    const uint64_t x12 = m61_mul(sq(x6, 6), x6);
    const uint64_t x24 = m61_mul(sq(x12, 12), x12);
    const uint64_t x48 = m61_mul(sq(x24, 24), x24);
    const uint64_t x56 = m61_mul(sq(x48, 8), (m61_mul(sq(x6, 2), x2)));  // x_8 = x_6^4 x_2
    const uint64_t x58 = m61_mul(sq(x56, 2), x2);
    const uint64_t x59 = m61_mul(sq(x58, 1), x1);
    return m61_mul(sq(x59, 2), x1);
}

struct FieldU64 {
    using T = uint64_t;
    uint64_t p;
    uint64_t pneg;  // -p^{-1} mod 2^64
    uint64_t r2;    // 2^128 mod p
    uint64_t m61;   // 1 when p = 2^61 - 1 (Mersenne reduction), else 0

    __host__ __device__ __forceinline__ uint64_t mont(uint64_t a, uint64_t b) const {
        const uint64_t t_lo = a * b;
        const uint64_t t_hi = mulhi64(a, b);
        if (m61) return m61_redc(t_lo, t_hi);
        const uint64_t m = t_lo * pneg;
        const uint64_t u = t_hi + mulhi64(m, p) + (t_lo != 0ull);
        return u >= p ? u - p : u;
    }
    __host__ __device__ __forceinline__ uint64_t sub(uint64_t a, uint64_t b) const {
        return a >= b ? a - b : a + (p - b);
    }
    __host__ __device__ __forceinline__ uint64_t add(uint64_t a, uint64_t b) const {
        return a >= p - b ? a - (p - b) : a + b;
    }
    __host__ __device__ __forceinline__ uint64_t to_mont(uint64_t a) const { return mont(a, r2); }
    // (h x - a y) R^{-1} mod p, the fraction-free row update of FieldU32 with
    // one shared reduction: t = h x + (p - a) y < 2 p^2 < p 2^64 (p < 2^63) is
    // a valid REDC input, summed in 128 bits.
    __host__ __device__ __forceinline__ void ff_sum(uint64_t h, uint64_t x, uint64_t a, uint64_t y, uint64_t& t_lo,
                                                    uint64_t& t_hi) const {
        const uint64_t b = p - a;
        const uint64_t l1 = h * x, l2 = b * y;
        t_lo = l1 + l2;
        t_hi = mulhi64(h, x) + mulhi64(b, y) + (t_lo < l1 ? 1ull : 0ull);
    }
    __host__ __device__ __forceinline__ uint64_t redc(uint64_t t_lo, uint64_t t_hi) const {
        const uint64_t m = t_lo * pneg;
        const uint64_t u = t_hi + mulhi64(m, p) + (t_lo != 0ull);
        return u >= p ? u - p : u;
    }
    __host__ __device__ __forceinline__ uint64_t ff_update(uint64_t h, uint64_t x, uint64_t a, uint64_t y) const {
        uint64_t t_lo, t_hi;
        ff_sum(h, x, a, y, t_lo, t_hi);
        if (m61) return m61_redc(t_lo, t_hi);
        return redc(t_lo, t_hi);
    }
    // N updates, the reduction chosen once for all of them
    template <int N>
    __host__ __device__ __forceinline__ void ff_update_n(uint64_t h, const uint64_t* x, const uint64_t* a, uint64_t y,
                                                         uint64_t* out) const {
        uint64_t lo[N], hi[N];
#pragma unroll
        for (int i = 0; i < N; ++i) ff_sum(h, x[i], a[i], y, lo[i], hi[i]);
        if (m61) {
#pragma unroll
            for (int i = 0; i < N; ++i) out[i] = m61_redc(lo[i], hi[i]);
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) out[i] = redc(lo[i], hi[i]);
        }
    }
    __host__ __device__ __forceinline__ uint64_t inverse(uint64_t h) const {
        const uint64_t e = p - 2ull;
        const uint64_t hm = to_mont(h);
        uint64_t r = to_mont(1ull);
        for (int bit = 63 - clz64(e); bit >= 0; --bit) {
            r = mont(r, r);
            if ((e >> bit) & 1ull) r = mont(r, hm);
        }
        return mont(r, 1ull);
    }
    __host__ __device__ __forceinline__ uint64_t factor_of_head(uint64_t h) const { return to_mont(to_mont(inverse(h))); }
    __host__ __device__ __forceinline__ uint64_t mul(uint64_t a, uint64_t b) const { return mont(to_mont(a), b); }
    __host__ __device__ __forceinline__ uint64_t reduce(uint64_t a) const { return a >= p ? a % p : a; }
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
    f.m61 = p == M61 ? 1ull : 0ull;
    return f;
}

}  // namespace linac
