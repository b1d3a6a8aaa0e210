// The float arctangent of the literal Heaviside (CUDA_RSF/kernel.cu:258),
// shared by the kernels and by the host build of its exhaustive check.
//
// The oracle's arctangent is S. L. Moshier's Cephes atan.c rational in double
// precision (two divisions), rounded once to float.  atanf_literal returns the
// same float for every float input at about half the instructions: a degree-9
// polynomial in double precision after a one-division range reduction.  Its
// double result lies within 2 units in the last place of the Cephes one, and
// rounded to float it equals the oracle's float on all 2^32 inputs (the host
// build of this header is checked against the oracle on every one of them).
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define RSF_HD __host__ __device__ __forceinline__
#else
#define RSF_HD inline
#endif
#ifdef __clang__
#define RSF_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define RSF_NO_CONTRACT
#endif

namespace rsf {

// Double division without range scaling.  The compiler's IEEE f64 division is
// div_scale(d), rcp, two Newton steps e = fma(-d, r, 1), r = fma(r, e, r),
// div_scale(n), q = n*r, a remainder correction fma(fma(-d, q, n), r, q)
// (div_fmas) and div_fixup.  For 0 < d in [2^-900, 2^900] and n a float value
// (or zero) the scaling steps return their inputs and div_fmas is a plain
// multiply-add, and the fixup's only effect is the sign of a zero quotient,
// which is that of n.  recip_nr and div_by run the remaining steps, so
// div_by(n, d, recip_nr(d)) is bit for bit `n / d` -- which the host build
// evaluates -- and one reciprocal serves several numerators.
RSF_HD double recip_nr(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r0 = __builtin_amdgcn_rcp(d);
    const double r1 = __builtin_fma(r0, __builtin_fma(-d, r0, 1.0), r0);
    return __builtin_fma(r1, __builtin_fma(-d, r1, 1.0), r1);
#else
    return 1.0 / d;   // unused by div_by on the host
#endif
}
RSF_HD double div_by(double n, double d, double r) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double q = n * r;
    return __builtin_copysign(__builtin_fma(__builtin_fma(-d, q, n), r, q), n);
#else
    (void)r;
    return n / d;
#endif
}

// atan_poly(ax) ~ atan(ax) for ax >= 0.  Reduction: ax <= tan(pi/8): v = ax;
// up to tan(3pi/8): v = (ax-1)/(ax+1) and pi/4; beyond: v = -1/ax and pi/2
// (ax clamped to 2^100, so infinity gives pi/2), so |v| <= tan(pi/8).  Then
// atan(v) = v + v*u*p(u), u = v^2, with p the degree-9 Chebyshev interpolant
// of (atan(sqrt u)/sqrt u - 1)/u on [0, 0.1716] (relative error 2^-52.7).
RSF_HD double atan_poly(double ax) {
    RSF_NO_CONTRACT
    const bool hi = ax > 2.41421356237309504880;               // tan(3 pi / 8)
    const bool mi = !hi && ax > 0.41421356237309504880;        // tan(pi / 8)
    const double num = hi ? -1.0 : (mi ? ax - 1.0 : ax);       // exact
    const double den = hi ? (ax < 0x1p100 ? ax : 0x1p100) : (mi ? ax + 1.0 : 1.0);
    const double v = div_by(num, den, recip_nr(den));
    const double base = hi ? 1.57079632679489661923 : (mi ? 7.85398163397448309616e-1 : 0.0);
    const double u = v * v;
    double p = 0x1.74b41a052a60ap-6;
    p = __builtin_fma(p, u, -0x1.6f427ef5b98a3p-5);
    p = __builtin_fma(p, u, 0x1.d5ea6ffbc6ff9p-5);
    p = __builtin_fma(p, u, -0x1.105e2879e082ep-4);
    p = __builtin_fma(p, u, 0x1.3b068cadba7f1p-4);
    p = __builtin_fma(p, u, -0x1.745c7f6370d7ap-4);
    p = __builtin_fma(p, u, 0x1.c71c6dd0c74efp-4);
    p = __builtin_fma(p, u, -0x1.2492491dd2b46p-3);
    p = __builtin_fma(p, u, 0x1.9999999990a6fp-3);
    p = __builtin_fma(p, u, -0x1.5555555555546p-2);
    return base + __builtin_fma(v * u, p, v);
}
// The literal path's float arctangent (NaN gives NaN; -0 gives +0, as the
// oracle's sign rule x < 0 ? -y : y does).
RSF_HD float atanf_literal(float xf) {
    const double x = (double)xf;
    const float f = (float)atan_poly(x < 0 ? -x : x);
    return x < 0 ? -f : f;
}

}  // namespace rsf
