// Real and complex float64 element operations shared by the float RREF paths.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace linac {

struct CplxF {
    double re, im;
};

template <class T>
struct FloatOps;

template <>
struct FloatOps<double> {
    static constexpr int kReal = 1;  // doubles per element
    __device__ static double abs(double x) { return fabs(x); }
    __device__ static double div(double a, double b) { return a / b; }
    __device__ static double fms(double x, double f, double y) { return fma(-f, y, x); }  // x - f*y
    __device__ static double scale(double x, double s) { return x / s; }
    __device__ static double zero() { return 0.0; }
    __device__ static double one() { return 1.0; }
    __device__ static bool is_zero(double x) { return x == 0.0; }
    __device__ static double mul(double a, double b) { return a * b; }
    __device__ static double neg(double a) { return -a; }
    __device__ static double sel(bool c, double a, double b) { return c ? a : b; }
    __device__ static double inv(double a) { return 1.0 / a; }
};

template <>
struct FloatOps<CplxF> {
    static constexpr int kReal = 2;
    __device__ static double abs(CplxF z) { return hypot(z.re, z.im); }
    __device__ static CplxF div(CplxF a, CplxF b) {
        // Smith's algorithm: no overflow for large |b|
        if (fabs(b.re) >= fabs(b.im)) {
            const double r = b.im / b.re, d = b.re + b.im * r;
            return CplxF{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
        }
        const double r = b.re / b.im, d = b.im + b.re * r;
        return CplxF{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
    }
    __device__ static CplxF fms(CplxF x, CplxF f, CplxF y) {
        return CplxF{x.re - (f.re * y.re - f.im * y.im), x.im - (f.re * y.im + f.im * y.re)};
    }
    __device__ static CplxF scale(CplxF x, double s) { return CplxF{x.re / s, x.im / s}; }
    __device__ static CplxF zero() { return CplxF{0.0, 0.0}; }
    __device__ static CplxF one() { return CplxF{1.0, 0.0}; }
    __device__ static bool is_zero(CplxF z) { return z.re == 0.0 && z.im == 0.0; }
    __device__ static CplxF mul(CplxF a, CplxF b) { return CplxF{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
    __device__ static CplxF inv(CplxF a) { return div(one(), a); }
    __device__ static CplxF neg(CplxF a) { return CplxF{-a.re, -a.im}; }
    // per component: a select of the whole struct is lowered through scratch memory
    __device__ static CplxF sel(bool c, CplxF a, CplxF b) { return CplxF{c ? a.re : b.re, c ? a.im : b.im}; }
};

}  // namespace linac
