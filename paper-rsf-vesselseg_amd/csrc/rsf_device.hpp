// Device-side helpers shared by the RSF kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "rsf_atan.hpp"

namespace rsf {

constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxTaps = 63;

// Taps travel as a by-value kernel argument, so they live in SGPRs.  g[0..P)
// are the taps; g[P..2P-1) repeat them (g[P + i] = g[i]), so that any P
// consecutive entries starting below P are the taps rotated.
struct Taps {
    float g[2 * kMaxTaps];
};

// One sub-volume of a (possibly batched) run: first voxel and dims.
struct BlockDesc {
    int64_t off;
    int w, h, l, pad_;
};

// A unit of work of the tiled sweeps: block index (< 0 = empty slot) and the
// tile's first voxel inside the block.
struct Tile {
    int b, x0, y0, z0;
};

// ---------------------------------------------------------------- point ops
// atanf_literal (the float arctangent of the literal path) and the unscaled
// double division recip_nr / div_by.
// Heaviside of CUDA_RSF/kernel.cu:258-259: float atan, double 2/pi mul-add.
// Contraction is off inside the literal helpers so that each operation is
// rounded on its own, as in the oracle.
__device__ __forceinline__ float heav_literal(float phi, float eps) {
#pragma clang fp contract(off)
    return (float)(0.5 * (1.0 + (2.0 / kPi) * (double)atanf_literal(phi / eps)));
}
// The literal Heaviside with phi / eps taken as phi * r when r = 1/eps is exact:
// for eps a power of two whose reciprocal is a normal float, both round the
// exact value phi * 2^-k once, so they give the same float for every phi
// (subnormal results and overflow included).  r = 0 keeps the division.
__device__ __forceinline__ float heav_literal_r(float phi, float eps, float r) {
#pragma clang fp contract(off)
    const float q = r != 0.0f ? phi * r : phi / eps;
    return (float)(0.5 * (1.0 + (2.0 / kPi) * (double)atanf_literal(q)));
}
// Dirac of kernel.cu:275-276: float quotient scaled by the double 1/pi.
__device__ __forceinline__ float dirac_literal(float phi, float eps) {
#pragma clang fp contract(off)
    return (float)((1.0 / kPi) * (double)(eps / ((eps * eps) + (phi * phi))));
}
// Square root without range scaling.  The compiler's IEEE sqrtf is: inputs
// below 2^-96 scaled by 2^32, the hardware root (within one unit), a step to
// the neighbour s -+ 1 unit whose residual fma(-s', s, x) changes sign, the
// scale undone, and +0 and +inf passed through.  sqrt_unscaled runs the root
// and the step only; for x >= 2^-96, +0, +inf and NaN no scaling applies and
// the step returns zero, infinity and NaN unchanged, so it is bit for bit
// sqrtf there.
__device__ __forceinline__ float sqrt_unscaled(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sd = __int_as_float(__float_as_int(s) - 1), su = __int_as_float(__float_as_int(s) + 1);
    const float r = __builtin_fmaf(-sd, s, x) <= 0.0f ? sd : s;
    return __builtin_fmaf(-su, s, x) > 0.0f ? su : r;
}
// Literal unit-normal components (float)(c / ((double)M + 1e-6)) with
// M = sqrtf(m2) the float |grad| (kernel.cu:185).  The exact form calls sqrtf
// and the compiler's division.  The fast form uses sqrt_unscaled and one
// reciprocal of the denominator for all components (div_by); it is bit for
// bit the exact form unless normal_outside(m2): m2 in (0, 2^-96) (sqrt
// scaling) or not finite (the denominator leaves div_by's range).  Finite M
// keeps the denominator in [1e-6, 2^128] and each c is a float.
__device__ __forceinline__ bool normal_outside(float m2) {
    const uint32_t b = __float_as_uint(m2);
    return b != 0u && b - 0x0f800000u > 0x7f7fffffu - 0x0f800000u;
}
template <int N>
__device__ __forceinline__ void normal_literal_fast(const float (&c)[N], float m2, float (&out)[N]) {
    const double d = (double)sqrt_unscaled(m2) + 1e-6;
    const double r = recip_nr(d);
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = (float)div_by((double)c[i], d, r);
}
template <int N>
__device__ __forceinline__ void normal_literal_exact(const float (&c)[N], float m2, float (&out)[N]) {
    const double d = (double)sqrtf(m2) + 1e-6;
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = (float)((double)c[i] / d);
}
// Centred Heaviside G = H - 1/2 = atan(phi/eps)/pi (FP32).
__device__ __forceinline__ float heav_centred(float phi, float inv_eps) {
    return atanf(phi * inv_eps) * (float)(1.0 / kPi);
}
__device__ __forceinline__ float dirac_fast(float phi, float eps) {
    return (float)(1.0 / kPi) * (eps * __builtin_amdgcn_rcpf(__builtin_fmaf(phi, phi, eps * eps)));
}

// Local fits of the fast mode from bI = blur(I), bIG = blur(I*G), bG = blur(G):
// blur(H) = (sum g)^3/2 + bG and blur(I*H) = bI/2 + bIG (H = 1/2 + G), likewise
// for 1-H.  Written out (no implicit contraction) so that every sweep-1 kernel
// rounds them identically.
__device__ __forceinline__ void fast_fits(float bI, float bIG, float bG, float halfS3, float& fo, float& fi,
                                          float& bH, float& bHc) {
#pragma clang fp contract(off)
    bH = halfS3 + bG;
    bHc = halfS3 - bG;
    fo = __builtin_fmaf(0.5f, bI, bIG) * __builtin_amdgcn_rcpf(bH);
    fi = __builtin_fmaf(0.5f, bI, -bIG) * __builtin_amdgcn_rcpf(bHc);
}
// F = lambda1*Eo - lambda2*Ei, Eo = (I-fo)^2, Ei = (I-fi)^2 (sigma2 = 0,
// kernel.cu:44-45; lambda1 == lambda2 gives e*(Eo-Ei) of kernel.cu:384).
__device__ __forceinline__ float force_of(float I, float fo, float fi, float l1, float l2, int same) {
#pragma clang fp contract(off)
    const float d1 = I - fo, d2 = I - fi;
    const float Eo = d1 * d1, Ei = d2 * d2;
    return same ? l1 * (Eo - Ei) : (l1 * Eo - l2 * Ei);
}

// The sweep-2 kernels write their fused multiply-adds out explicitly (and
// contract nothing else in these helpers), so that every kernel variant
// rounds the update identically whichever of them a voxel runs on.
// (a + b + c)/2 in the reference's order ((a/2 + b/2) + c/2); the halvings are
// exact, so the two fused forms round exactly like the separate operations.
__device__ __forceinline__ float half_sum3(float a, float b, float c) {
    return __builtin_fmaf(c, 0.5f, __builtin_fmaf(a, 0.5f, b * 0.5f));
}
// |g|^2 = gx^2 + gy^2 + gz^2 (kernel.cu:914-933).  The literal path rounds each
// product and sum on its own, left to right, as the reference kernel reads; the
// fast path fuses two multiply-adds.
template <bool LIT>
__device__ __forceinline__ float mag2(float gx, float gy, float gz) {
#pragma clang fp contract(off)
    if constexpr (LIT) return (gx * gx + gy * gy) + gz * gz;
    else return __builtin_fmaf(gz, gz, __builtin_fmaf(gy, gy, gx * gx));
}
// phi' = ((phi - (delta*F)*dt) + ((nu*dt)*delta)*kappa) + mu*(dt*(L - kappa))
// (kernel.cu:384-386, F = lambda1*Eo - lambda2*Ei); nudt = nu*dt.  Literal:
// every operation rounded on its own; fast: three fused multiply-adds.
template <bool LIT>
__device__ __forceinline__ float phi_update(float p0, float d, float F, float kap, float lap, float dt, float nudt,
                                            float mu) {
#pragma clang fp contract(off)
    if constexpr (LIT) {
        return ((p0 - (d * F) * dt) + (nudt * d) * kap) + mu * (dt * (lap - kap));
    } else {
        const float a = __builtin_fmaf(-(d * F), dt, p0);
        const float b = __builtin_fmaf(nudt * d, kap, a);
        return __builtin_fmaf(mu, dt * (lap - kap), b);
    }
}

// ------------------------------------------------- difference operator rules
// D f(q) = (f(R) - f(L)) / 2 with the reference edge rules (kernel.cu:57-125):
//   x: R = q+1 (q at x = w-1), L = q-1 in linear order (wraps to the previous
//      row at x = 0; element 0 at linear index 0);
//   y, z: interior central, one-sided halves at both ends.
struct Grad {
    float x, y, z;
};

__device__ __forceinline__ Grad grad_at(const float* __restrict__ f, int qx, int qy, int qz,
                                        int w, int h, int l) {
    const int64_t plane = (int64_t)w * h;
    const int64_t qi = (int64_t)qz * plane + (int64_t)qy * w + qx;
    const int64_t xr = (qx == w - 1) ? qi : qi + 1;
    const int64_t xl = (qi == 0) ? 0 : qi - 1;
    int64_t yr, yl, zr, zl;
    if (qy == 0) { yr = qi + w; yl = qi; }
    else if (qy == h - 1) { yr = qi; yl = qi - w; }
    else { yr = qi + w; yl = qi - w; }
    if (qz == 0) { zr = qi + plane; zl = qi; }
    else if (qz == l - 1) { zr = qi; zl = qi - plane; }
    else { zr = qi + plane; zl = qi - plane; }
    Grad g;
    g.x = (f[xr] - f[xl]) * 0.5f;
    g.y = (f[yr] - f[yl]) * 0.5f;
    g.z = (f[zr] - f[zl]) * 0.5f;
    return g;
}

// Coordinates of the D_x left operand of voxel (x,y,z): the previous linear
// element.
__device__ __forceinline__ void left_x(int x, int y, int z, int w, int h, int& qx, int& qy,
                                       int& qz) {
    if (x > 0) { qx = x - 1; qy = y; qz = z; }
    else if (y > 0) { qx = w - 1; qy = y - 1; qz = z; }
    else if (z > 0) { qx = w - 1; qy = h - 1; qz = z - 1; }
    else { qx = 0; qy = 0; qz = 0; }
}

}  // namespace rsf
