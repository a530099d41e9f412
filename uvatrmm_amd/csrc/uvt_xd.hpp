// uvt_xd.hpp -- FP64 numbers with a separate int exponent, for norms and
// error bounds whose terms may lie anywhere in (or beyond) the FP64 range.
//
// An xd is f * 2^e with f == 0 or 0.5 <= |f| < 1, so products and sums of
// terms like 2^-alpha_i * ||A_i|| never overflow or underflow; a sum drops
// a term only when it is below 2^-1074 of the larger one.  Scaling by a power
// of two is exact.
//
// Sums of squares and of absolute values over input entries use a cheap
// split: entries with 2^-400 <= |x| <= 2^400 (or 0) are summed in plain FP64,
// where x^2 times any depth this engine allows stays inside the normal range;
// the rare others go to an xd accumulator, and non-finite entries set a flag.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace uvt {

struct xd {
    double f;
    int e;
};

__host__ __device__ __forceinline__ xd xd_norm(double f, int e)
{
    if (f == 0.0) return xd{0.0, 0};
    int k = 0;
    const double g = frexp(f, &k);
    return xd{g, e + k};
}
__host__ __device__ __forceinline__ xd xd_from(double v) { return xd_norm(v, 0); }
__host__ __device__ __forceinline__ xd xd_pow2(int k) { return xd{0.5, k + 1}; }
__host__ __device__ __forceinline__ xd xd_mul(xd a, xd b) { return xd_norm(a.f * b.f, a.e + b.e); }
__host__ __device__ __forceinline__ xd xd_add(xd a, xd b)
{
    if (a.f == 0.0) return b;
    if (b.f == 0.0) return a;
    const int e = a.e > b.e ? a.e : b.e;
    // the smaller term's shift is clamped: below 2^-1100 of the larger it is 0
    const int da = a.e - e < -1100 ? -1100 : a.e - e, db = b.e - e < -1100 ? -1100 : b.e - e;
    return xd_norm(ldexp(a.f, da) + ldexp(b.f, db), e);
}
__host__ __device__ __forceinline__ xd xd_div(xd a, xd b) { return xd_norm(a.f / b.f, a.e - b.e); }
__host__ __device__ __forceinline__ xd xd_sqrt(xd a)
{
    if (a.f == 0.0) return a;
    const bool odd = (a.e & 1) != 0;
    const double f = odd ? a.f * 2.0 : a.f;
    const int e = odd ? a.e - 1 : a.e;  // even
    return xd_norm(sqrt(f), e / 2);
}
// The value as an FP64 number (overflows to inf, underflows to 0).
__host__ __device__ __forceinline__ double xd_val(xd a)
{
    const int e = a.e > 2000 ? 2000 : (a.e < -2000 ? -2000 : a.e);
    return ldexp(a.f, e);
}
__host__ __device__ __forceinline__ bool xd_lt(xd a, xd b)  // a < b for a, b >= 0
{
    if (a.f == 0.0) return b.f > 0.0;
    if (b.f == 0.0) return false;
    return a.e != b.e ? a.e < b.e : a.f < b.f;
}

__device__ __forceinline__ xd xd_shfl_xor(xd a, int off)
{
    return xd{__shfl_xor(a.f, off, 64), __shfl_xor(a.e, off, 64)};
}
// Sum over the 64 lanes of a wave (every lane gets the total).
__device__ __forceinline__ xd xd_wave_sum(xd a)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = xd_add(a, xd_shfl_xor(a, off));
    return a;
}

// Sum of x^2 and of |x| over a stream of entries (see the file comment).
struct NormAcc {
    double s2 = 0.0, s1 = 0.0;  // entries in [2^-400, 2^400] or 0
    xd x2{0.0, 0}, x1{0.0, 0};  // the others
    bool nonfinite = false;

    __device__ __forceinline__ void add(double x)
    {
        const double ax = fabs(x);
        const bool mid = (ax <= 0x1p400) && (ax >= 0x1p-400 || ax == 0.0);
        if (!mid) {  // rare: almost never taken by any lane of a wave
            if (!(ax <= 1.79769313486231570815e308)) {
                nonfinite = true;
            } else {
                const xd a = xd_from(ax);
                x2 = xd_add(x2, xd_mul(a, a));
                x1 = xd_add(x1, a);
            }
        }
        const double xm = mid ? ax : 0.0;
        s2 = fma(xm, xm, s2);
        s1 += xm;
    }
    // Totals over the wave (every lane gets them).  The xd parts are summed
    // only when some lane of the wave holds one.
    __device__ __forceinline__ void wave_reduce()
    {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            s2 += __shfl_xor(s2, off, 64);
            s1 += __shfl_xor(s1, off, 64);
        }
        nonfinite = __any((int)nonfinite) != 0;
        if (__any(x2.f != 0.0 || x1.f != 0.0)) {
            x2 = xd_wave_sum(x2);
            x1 = xd_wave_sum(x1);
        }
    }
    __device__ __forceinline__ xd sum_sq() const { return xd_add(xd_from(s2), x2); }
    __device__ __forceinline__ xd sum_abs() const { return xd_add(xd_from(s1), x1); }
};

}  // namespace uvt
