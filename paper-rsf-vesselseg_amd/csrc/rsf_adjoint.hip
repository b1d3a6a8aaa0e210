// Adjoint (vector-Jacobian product) of one RSF step on the device.
//
// Forward step (CUDA_RSF/main.cpp:234-300, sigma2 = 0):
//   phi' = phi - dt*delta*F + nu*dt*delta*kappa + mu*dt*(L - kappa)
//   F = lambda1*(I-fo)^2 - lambda2*(I-fi)^2, fo = PA/PH, fi = PB/PHc (valid
//   region; the pad band has fo = fi = 1), PA = blur(I*H), PH = blur(H),
//   PB = blur(I*(1-H)), PHc = blur(1-H), H' = delta,
//   kappa = sum_a D_a n_a, n = g/(|g|+1e-6), L = sum_a D_a g_a, g = D phi.
// Given u = dLoss/dphi':
//   Fbar = -dt*delta*u, fobar = -2(I-fo)*lambda1*Fbar, fibar = 2(I-fi)*lambda2*Fbar,
//   Hbar = blur^T(QH) + I*blur^T(QA) with QH = PHbar - PHcbar, QA = PAbar - PBbar,
//   nbar_a = D_a^T(kbar), kbar = (nu*dt*delta - mu*dt)*u, Lbar = mu*dt*u,
//   G_a = nbar_a/(M+1e-6) - g_a*(nbar.g)/(M(M+1e-6)^2) + D_a^T(Lbar),
//   dphi = u + delta'*(-dt*F + nu*dt*kappa)*u + delta*Hbar + sum_a D_a^T G_a.
// The taps are symmetric, so blur^T is the same correlation applied to the
// valid-region adjoint with zeros outside.  D^T follows the reference edge
// rules: along x the left operand is the previous linear element (wrap), so
// its transpose gathers from the next linear element.
#include <initializer_list>

#include "rsf_internal.hpp"

namespace rsf {
namespace adj {
namespace {

constexpr int kT = 256;
constexpr float kEta = 1e-6f;

inline dim3 grid_for(int64_t n) {
    int64_t b = (n + kT - 1) / kT;
    if (b > (1 << 16)) b = 1 << 16;
    return dim3((unsigned)(b < 1 ? 1 : b));
}

#define VOXELS(i, n) \
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// x, y, z of a block-local linear index; 32-bit division when the block allows it
__device__ __forceinline__ void coords(int64_t i, int w, int h, int64_t n, int& x, int& y, int& z) {
    if (n < (int64_t(1) << 31)) {
        const uint32_t ii = (uint32_t)i, r = ii / (uint32_t)w;
        x = (int)(ii - r * (uint32_t)w);
        z = (int)(r / (uint32_t)h);
        y = (int)(r - (uint32_t)z * (uint32_t)h);
    } else {
        const int64_t r = i / w;
        x = (int)(i - r * w);
        z = (int)(r / h);
        y = (int)(r - (int64_t)z * h);
    }
}

__device__ __forceinline__ float dirac(float p, float eps) {
#pragma clang fp contract(off)
    return (float)(1.0 / kPi) * eps / (eps * eps + p * p);
}
__device__ __forceinline__ float dirac_prime(float p, float eps) {
#pragma clang fp contract(off)
    const float q = eps * eps + p * p;
    return -(float)(2.0 / kPi) * eps * p / (q * q);
}
__device__ __forceinline__ float image_at(const void* img, int u8, int64_t i) {
    return u8 ? (float)((const uint8_t*)img)[i] : ((const float*)img)[i];
}

// (D_a^T u)(j) for a value getter u(q) over block-local linear indices.
template <class U>
__device__ __forceinline__ float dT_x(const U& u, int64_t j, int x, int w, int64_t N) {
#pragma clang fp contract(off)
    float plus = 0.f, minus = 0.f;
    if (x != 0) plus += u(j - 1);
    if (x == w - 1) plus += u(j);
    if (j + 1 < N) minus += u(j + 1);
    if (j == 0) minus += u(0);
    return (plus - minus) * 0.5f;
}
template <class U>
__device__ __forceinline__ float dT_yz(const U& u, int64_t j, int c, int n, int64_t stride) {
#pragma clang fp contract(off)
    float plus = 0.f, minus = 0.f;
    if (c >= 1) plus += u(j - stride);
    if (c == n - 1) plus += u(j);
    if (c == 0) minus += u(j);
    if (c <= n - 2) minus += u(j + stride);
    return (plus - minus) * 0.5f;
}

// The co-states of the blurred channels at voxel i: QH = PHbar - PHcbar and QA =
// PAbar - PBbar inside the valid region, 0 in the pad band.
struct QArgs {
    const float *phi, *gout, *fo, *fi, *PH, *PHc;
    const void* img;
    int u8, m;
    float dt, l1, l2, eps;
};
__device__ __forceinline__ void q_at(const QArgs& a, int64_t i, int x, int y, int z, int w, int h, int l, float& qh,
                                     float& qa) {
    const int m = a.m;
    const bool valid = x >= m && x < w - m && y >= m && y < h - m && z >= m && z < l - m;
    qh = 0.f;
    qa = 0.f;
    if (valid) {
        const float I = image_at(a.img, a.u8, i);
        const float Fb = -a.dt * dirac(a.phi[i], a.eps) * a.gout[i];
        const float fob = -2.f * (I - a.fo[i]) * a.l1 * Fb;
        const float fib = 2.f * (I - a.fi[i]) * a.l2 * Fb;
        const float ph = a.PH[i], phc = a.PHc[i];
        qh = -fob * a.fo[i] / ph + fib * a.fi[i] / phc;   // PHbar - PHcbar
        qa = fob / ph - fib / phc;                        // PAbar - PBbar
    }
}

__device__ __forceinline__ QArgs shift_q(QArgs a, int64_t o) {
    a.phi += o;
    a.gout += o;
    a.fo += o;
    a.fi += o;
    a.PH += o;
    a.PHc += o;
    a.img = static_cast<const char*>(a.img) + o * (a.u8 ? 1 : 4);
    return a;
}

__global__ void k_pointwise_q(float* QH, float* QA, QArgs a, int w, int h, int l) {
    const int64_t plane = (int64_t)w * h, n = plane * l;
    VOXELS(i, n) {
        int x, y, z;
        coords(i, w, h, n, x, y, z);
        float qh, qa;
        q_at(a, i, x, y, z, w, h, l, qh, qa);
        QH[i] = qh;
        QA[i] = qa;
    }
}

// The co-states and the x pass of their transposed blur in one kernel: a 64 x 4
// thread block takes 4 rows at a time, forms QH and QA of the rows in LDS (2 x 4 x w
// floats of dynamic LDS), and correlates them along x from there, every lane
// running the same masked taps as k_conv_same_zero2 along x (the same bits as the
// two kernels, without QH and QA through HBM).  With a block table bt, blockIdx.y
// is the block (every block of the launch w x h x l), its first voxel bt[.].off.
template <int P>
__global__ __launch_bounds__(256) void k_q_convx(float* __restrict__ out0, float* __restrict__ out1, QArgs a, int w,
                                                 int h, int l, Taps t, const BlockDesc* __restrict__ bt) {
    constexpr int m = (P - 1) / 2;
    if (bt) {
        const int64_t o = bt[blockIdx.y].off;
        out0 += o;
        out1 += o;
        a = shift_q(a, o);
    }
    extern __shared__ float sq[];   // [4][2][w]
    float* row = sq + threadIdx.y * 2 * w;
    const int rows = h * l;
    // every block runs the same number of iterations (the barriers): rows past the
    // end compute nothing and store nothing
    for (int r0 = blockIdx.x * 4; r0 < rows; r0 += gridDim.x * 4) {
        const int r = r0 + threadIdx.y;
        const bool live = r < rows;
        const int y = live ? r % h : 0, z = live ? r / h : 0;
        const int64_t base = (int64_t)r * w;
        if (live) {
            for (int x = threadIdx.x; x < w; x += 64) {
                float qh, qa;
                q_at(a, base + x, x, y, z, w, h, l, qh, qa);
                row[x] = qh;
                row[w + x] = qa;
            }
        }
        __syncthreads();
        if (live) {
            for (int x = threadIdx.x; x < w; x += 64) {
                float a0 = 0.f, a1 = 0.f;
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    const int cc = x - m + k;
                    const bool in_range = cc >= 0 && cc < w;
                    const int q = min(max(cc, 0), w - 1);
                    a0 = fmaf(t.g[k], in_range ? row[q] : 0.f, a0);
                    a1 = fmaf(t.g[k], in_range ? row[w + q] : 0.f, a1);
                }
                out0[base + x] = a0;
                out1[base + x] = a1;
            }
        }
        __syncthreads();
    }
}

// out(x) = sum_k g[k] * in(x - m + k) along one axis, in = 0 outside the block.
__global__ void k_conv_same_zero(float* out, const float* in, int w, int h, int l, int axis, Taps t, int P, int m) {
    const int64_t plane = (int64_t)w * h, n = plane * l;
    const int dim = axis == 0 ? w : (axis == 1 ? h : l);
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)w : plane);
    VOXELS(i, n) {
        const int c = axis == 0 ? (int)(i % w) : (axis == 1 ? (int)((i / w) % h) : (int)(i / plane));
        float acc = 0.f;
        for (int k = 0; k < P; ++k) {
            const int cc = c - m + k;
            if (cc >= 0 && cc < dim) acc = fmaf(t.g[k], in[i + (int64_t)(cc - c) * stride], acc);
        }
        out[i] = acc;
    }
}

// The same correlation for two channels at once, with the axis and the tap count
// fixed at compile time (the x pass: consecutive lanes read consecutive
// columns, which a thread's own run of columns would not): a 64 x 4 thread
// block covers 64 x-voxels of 4 rows
// (y + z*h), rows grid-strided, no 64-bit division; along y and z interior rows
// take the unrolled taps without bounds checks, along x every lane runs the
// same masked taps.  The taps are summed in the order of k_conv_same_zero.
template <int AXIS, int P>
__global__ __launch_bounds__(256) void k_conv_same_zero2(float* __restrict__ out0, float* __restrict__ out1,
                                                         const float* __restrict__ in0,
                                                         const float* __restrict__ in1, int w, int h, int l, Taps t) {
    constexpr int m = (P - 1) / 2;
    const int x = blockIdx.x * 64 + threadIdx.x;
    if (x >= w) return;
    const int rows = h * l;
    const int dim = AXIS == 0 ? w : (AXIS == 1 ? h : l);
    const int64_t stride = AXIS == 0 ? 1 : (AXIS == 1 ? (int64_t)w : (int64_t)w * h);
    for (int r = blockIdx.y * 4 + threadIdx.y; r < rows; r += gridDim.y * 4) {
        const int y = r % h, z = r / h;
        const int c = AXIS == 0 ? x : (AXIS == 1 ? y : z);
        const int64_t i = (int64_t)r * w + x;
        float a0 = 0.f, a1 = 0.f;
        if constexpr (AXIS == 0) {
            // along x the lanes of a wave straddle the block's ends: no branch, the
            // out-of-range taps read a clamped neighbour and add g * 0
#pragma unroll
            for (int k = 0; k < P; ++k) {
                const int cc = c - m + k;
                const bool in_range = cc >= 0 && cc < dim;
                const int64_t q = i + (int64_t)(min(max(cc, 0), dim - 1) - c);
                a0 = fmaf(t.g[k], in_range ? in0[q] : 0.f, a0);
                a1 = fmaf(t.g[k], in_range ? in1[q] : 0.f, a1);
            }
        } else if (c >= m && c + m < dim) {
            const float* p0 = in0 + i - m * stride;
            const float* p1 = in1 + i - m * stride;
#pragma unroll
            for (int k = 0; k < P; ++k) {
                a0 = fmaf(t.g[k], p0[k * stride], a0);
                a1 = fmaf(t.g[k], p1[k * stride], a1);
            }
        } else {
            for (int k = 0; k < P; ++k) {
                const int cc = c - m + k;
                if (cc >= 0 && cc < dim) {
                    a0 = fmaf(t.g[k], in0[i + (int64_t)(cc - c) * stride], a0);
                    a1 = fmaf(t.g[k], in1[i + (int64_t)(cc - c) * stride], a1);
                }
            }
        }
        out0[i] = a0;
        out1[i] = a1;
    }
}

// The y and z passes of k_conv_same_zero2 with R consecutive outputs along the
// axis per thread: the R + P - 1 inputs of a thread's outputs are loaded once
// (R + P - 1 loads for R outputs instead of R P; y 22.6 -> 10.9 us, z 20.0 -> 10.8
// us for two 128^3 channels).  Out-of-range taps are skipped by selection, so
// every output is the same chain as k_conv_same_zero2's.  A 64 x 4 thread block:
// 64 x-voxels of 4 work items (the other coordinate and the group of R along the
// axis), work items grid-strided.  (Along x, R consecutive columns per thread made
// the pass slower, 20 -> 43 us: a wave's loads then spread over 2 KB each.)
template <int AXIS, int P, int R>
__global__ __launch_bounds__(256) void k_conv_same_zero2r(float* __restrict__ out0, float* __restrict__ out1,
                                                          const float* __restrict__ in0,
                                                          const float* __restrict__ in1, int w, int h, int l,
                                                          Taps t, const BlockDesc* __restrict__ bt) {
    static_assert(AXIS == 1 || AXIS == 2, "y or z");
    if (bt) {   // blockIdx.z: the block, as in k_q_convx
        const int64_t o = bt[blockIdx.z].off;
        out0 += o;
        out1 += o;
        in0 += o;
        in1 += o;
    }
    constexpr int m = (P - 1) / 2, NV = R + P - 1;
    const int x = blockIdx.x * 64 + threadIdx.x;
    if (x >= w) return;
    const int dim = AXIS == 1 ? h : l, other = AXIS == 1 ? l : h;
    const int groups = (dim + R - 1) / R;
    const int64_t plane = (int64_t)w * h;
    const int64_t stride = AXIS == 1 ? (int64_t)w : plane;
    const int items = groups * other;
    for (int it = blockIdx.y * 4 + threadIdx.y; it < items; it += gridDim.y * 4) {
        const int o = it / groups, c0 = (it - o * groups) * R;
        const int64_t base = (AXIS == 1 ? (int64_t)o * plane : (int64_t)o * w) + x;   // c = 0 of the line
        float v0[NV], v1[NV];
        if (c0 >= m && c0 + R - 1 + m < dim) {
            const float* p0 = in0 + base + (int64_t)(c0 - m) * stride;
            const float* p1 = in1 + base + (int64_t)(c0 - m) * stride;
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                v0[k] = p0[k * stride];
                v1[k] = p1[k * stride];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float a0 = 0.f, a1 = 0.f;
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    a0 = fmaf(t.g[k], v0[r + k], a0);
                    a1 = fmaf(t.g[k], v1[r + k], a1);
                }
                out0[base + (int64_t)(c0 + r) * stride] = a0;
                out1[base + (int64_t)(c0 + r) * stride] = a1;
            }
        } else {
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const int cc = min(max(c0 - m + k, 0), dim - 1);
                v0[k] = in0[base + (int64_t)cc * stride];
                v1[k] = in1[base + (int64_t)cc * stride];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float a0 = 0.f, a1 = 0.f;
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    const int cc = c0 + r - m + k;
                    const bool in = cc >= 0 && cc < dim;
                    a0 = in ? fmaf(t.g[k], v0[r + k], a0) : a0;
                    a1 = in ? fmaf(t.g[k], v1[r + k], a1) : a1;
                }
                if (c0 + r < dim) {
                    out0[base + (int64_t)(c0 + r) * stride] = a0;
                    out1[base + (int64_t)(c0 + r) * stride] = a1;
                }
            }
        }
    }
}

// the y and z passes: Bh, Ba (the x pass) -> t1, t2 -> Bh, Ba
template <int P>
int conv2_yz(float* Bh, float* Ba, float* t1, float* t2, int w, int h, int l, const Taps& taps, hipStream_t s,
             const BlockDesc* bt = nullptr, int nb = 1) {
    constexpr int R = 8;
    const dim3 blk(64, 4);
    const dim3 gy((unsigned)((w + 63) / 64), (unsigned)std::min<int64_t>(((int64_t)((h + R - 1) / R) * l + 3) / 4, 65535),
                  (unsigned)nb);
    hipLaunchKernelGGL((k_conv_same_zero2r<1, P, R>), gy, blk, 0, s, t1, t2, (const float*)Bh, (const float*)Ba, w, h,
                       l, taps, bt);
    const dim3 gz((unsigned)((w + 63) / 64), (unsigned)std::min<int64_t>(((int64_t)((l + R - 1) / R) * h + 3) / 4, 65535),
                  (unsigned)nb);
    hipLaunchKernelGGL((k_conv_same_zero2r<2, P, R>), gz, blk, 0, s, Bh, Ba, (const float*)t1, (const float*)t2, w, h,
                       l, taps, bt);
    RSF_HIP(hipGetLastError());
    return RSF_OK;
}

template <int P>
int conv2_axes(float* Bh, float* Ba, const float* QH, const float* QA, float* t1, float* t2, int w, int h, int l,
               const Taps& taps, hipStream_t s) {
    const dim3 blk(64, 4);
    const dim3 grid((unsigned)((w + 63) / 64), (unsigned)std::min<int64_t>(((int64_t)h * l + 3) / 4, 65535));
    hipLaunchKernelGGL((k_conv_same_zero2<0, P>), grid, blk, 0, s, Bh, Ba, QH, QA, w, h, l, taps);
    return conv2_yz<P>(Bh, Ba, t1, t2, w, h, l, taps, s);
}

// the co-states formed inside the x pass (k_q_convx), then the y and z passes
template <int P>
int q_conv2_axes(float* Bh, float* Ba, const QArgs& a, float* t1, float* t2, int w, int h, int l, const Taps& taps,
                 hipStream_t s, const BlockDesc* bt, int nb) {
    const dim3 blk(64, 4);
    const dim3 grid((unsigned)std::min<int64_t>(((int64_t)h * l + 3) / 4, 65535), (unsigned)nb);
    hipLaunchKernelGGL((k_q_convx<P>), grid, blk, (size_t)8 * w * sizeof(float), s, Bh, Ba, a, w, h, l, taps, bt);
    return conv2_yz<P>(Bh, Ba, t1, t2, w, h, l, taps, s, bt, nb);
}

// G_a at every voxel: normalisation Jacobian applied to D^T(kbar) plus D^T(Lbar).
// Also the forward normal and gradient at the voxel, for k_finish.
__global__ void k_stencil_g(float* Gx, float* Gy, float* Gz, FwdStencil fwd, const float* phi, const float* gout,
                            int w, int h, int l, float dt, float nu, float mu, float eps) {
#pragma clang fp contract(off)
    const int64_t plane = (int64_t)w * h, n = plane * l;
    auto kbar = [&](int64_t q) { return (nu * dt * dirac(phi[q], eps) - mu * dt) * gout[q]; };
    auto lbar = [&](int64_t q) { return mu * dt * gout[q]; };
    VOXELS(j, n) {
        int x, y, z;
        coords(j, w, h, n, x, y, z);
        const float nbx = dT_x(kbar, j, x, w, n), nby = dT_yz(kbar, j, y, h, w), nbz = dT_yz(kbar, j, z, l, plane);
        const float lbx = dT_x(lbar, j, x, w, n), lby = dT_yz(lbar, j, y, h, w), lbz = dT_yz(lbar, j, z, l, plane);
        const Grad g = grad_at(phi, x, y, z, w, h, l);
        const float M = sqrtf(g.x * g.x + g.y * g.y + g.z * g.z);
        const float r = 1.0f / (M + kEta);
        const float c = M > 0.f ? (nbx * g.x + nby * g.y + nbz * g.z) * r * r / M : 0.f;
        Gx[j] = nbx * r - g.x * c + lbx;
        Gy[j] = nby * r - g.y * c + lby;
        Gz[j] = nbz * r - g.z * c + lbz;
        fwd.n[0][j] = g.x / (M + kEta);
        fwd.n[1][j] = g.y / (M + kEta);
        fwd.n[2][j] = g.z / (M + kEta);
        fwd.g[0][j] = g.x;
        fwd.g[1][j] = g.y;
        fwd.g[2][j] = g.z;
    }
}

// Per-workgroup partial sums in a fixed order: a shuffle tree within each 64-lane
// wave, then the waves; row blockIdx.y * gridDim.x + blockIdx.x of partials.
__device__ __forceinline__ void block_partials(float* __restrict__ partials, float s0, float s1, float s2, float s3) {
    for (int o = 32; o > 0; o >>= 1) {
        s0 += __shfl_down(s0, o, 64);
        s1 += __shfl_down(s1, o, 64);
        s2 += __shfl_down(s2, o, 64);
        s3 += __shfl_down(s3, o, 64);
    }
    constexpr int kW = kT / 64;
    __shared__ float red[4][kW];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wv] = s0;
        red[1][wv] = s1;
        red[2][wv] = s2;
        red[3][wv] = s3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        float v = red[threadIdx.x][0];
        for (int q = 1; q < kW; ++q) v += red[threadIdx.x][q];
        partials[4 * (blockIdx.y * gridDim.x + blockIdx.x) + threadIdx.x] = v;
    }
}

// dphi and per-workgroup partial sums of the parameter gradients.
__global__ __launch_bounds__(kT) void k_finish(float* gin, float* partials, const float* phi, const float* gout,
                                               const float* Gx, const float* Gy, const float* Gz, FwdStencil fwd,
                                               const float* Bh,
                                               const float* Ba, const void* img, int u8, const float* fo,
                                               const float* fi, int w, int h, int l, float dt, float nu, float mu,
                                               float eps, float l1, float l2) {
#pragma clang fp contract(off)
    const int64_t plane = (int64_t)w * h, n = plane * l;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    auto gx = [&](int64_t q) { return Gx[q]; };
    auto gy = [&](int64_t q) { return Gy[q]; };
    auto gz = [&](int64_t q) { return Gz[q]; };
    VOXELS(i, n) {
        int x, y, z;
        coords(i, w, h, n, x, y, z);
        const float st = dT_x(gx, i, x, w, n) + dT_yz(gy, i, y, h, w) + dT_yz(gz, i, z, l, plane);
        // forward curvature and Laplacian at i (same stencil as sweep 2), from the
        // neighbours' normals and gradients (k_stencil_g): the right x neighbour (i
        // itself at x = w-1), the left one the previous linear element (wrap, element
        // 0 at i = 0), y and z one-sided at the ends
        const int64_t iRx = x == w - 1 ? i : i + 1, iLx = i == 0 ? 0 : i - 1;
        const int64_t iRy = y == h - 1 ? i : i + w, iLy = y == 0 ? i : i - w;
        const int64_t iRz = z == l - 1 ? i : i + plane, iLz = z == 0 ? i : i - plane;
        const float kap = ((fwd.n[0][iRx] - fwd.n[0][iLx]) + (fwd.n[1][iRy] - fwd.n[1][iLy]) +
                           (fwd.n[2][iRz] - fwd.n[2][iLz])) * 0.5f;
        const float lap = ((fwd.g[0][iRx] - fwd.g[0][iLx]) + (fwd.g[1][iRy] - fwd.g[1][iLy]) +
                           (fwd.g[2][iRz] - fwd.g[2][iLz])) * 0.5f;
        const float I = image_at(img, u8, i);
        const float d1 = I - fo[i], d2 = I - fi[i];
        const float Eo = d1 * d1, Ei = d2 * d2;
        const float F = l1 * Eo - l2 * Ei;
        const float p0 = phi[i], u = gout[i];
        const float d = dirac(p0, eps);
        const float Fb = -dt * d * u;
        gin[i] = u + dirac_prime(p0, eps) * (-dt * F + nu * dt * kap) * u + d * (Bh[i] + I * Ba[i]) + st;
        s0 += Eo * Fb;
        s1 -= Ei * Fb;
        s2 += dt * d * kap * u;
        s3 += dt * (lap - kap) * u;
    }
    block_partials(partials, s0, s1, s2, s3);
}

// k_stencil_g and k_finish on four consecutive x voxels per thread: the same
// arithmetic per voxel, on 16-byte loads and stores and 32-bit indices, the x
// neighbours shared within the quad.  For rows a multiple of 4 long, blocks under
// 2^31 voxels and arrays 16-byte aligned (a uint8 image 4-byte aligned).  The
// neighbour rows and planes are read clamped to the quad's own at the block's
// faces, where the one-sided rules take the voxel itself or drop the term.  Both
// kernel pairs compile without contraction, so that no multiply-add is fused in
// one and not the other (which the compiler decides by how values are shared): a
// block gives the same bits on either.
__device__ __forceinline__ void ld4(float (&a)[4], const float* __restrict__ p, uint32_t i) {
    const float4 v = *reinterpret_cast<const float4*>(p + i);
    a[0] = v.x;
    a[1] = v.y;
    a[2] = v.z;
    a[3] = v.w;
}
__device__ __forceinline__ void st4(float* __restrict__ p, uint32_t i, const float (&a)[4]) {
    *reinterpret_cast<float4*>(p + i) = make_float4(a[0], a[1], a[2], a[3]);
}
// a row's values at x0-1 .. x0+4: the previous linear element (element 0 for the
// block's first voxel), the quad, the next linear element (clamped at the end)
__device__ __forceinline__ void ld6(float (&a)[6], const float* __restrict__ p, uint32_t base, uint32_t n) {
    a[0] = p[base == 0 ? 0 : base - 1];
    a[5] = p[base + 4 < n ? base + 4 : n - 1];
    float c[4];
    ld4(c, p, base);
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i + 1] = c[i];
}
// dT_x and dT_yz on values: uL = u(j-1), u0 = u(j), uR = u(j+1) (uZ = u(0) at j = 0);
// um, up the neighbours along y or z
__device__ __forceinline__ float dTx_v(float uL, float u0, float uR, uint32_t j, int x, int w, uint32_t n) {
#pragma clang fp contract(off)
    float plus = 0.f, minus = 0.f;
    if (x != 0) plus += uL;
    if (x == w - 1) plus += u0;
    if (j + 1 < n) minus += uR;
    if (j == 0) minus += u0;
    return (plus - minus) * 0.5f;
}
__device__ __forceinline__ float dTyz_v(float um, float u0, float up, int c, int nn) {
#pragma clang fp contract(off)
    float plus = 0.f, minus = 0.f;
    if (c >= 1) plus += um;
    if (c == nn - 1) plus += u0;
    if (c == 0) minus += u0;
    if (c <= nn - 2) minus += up;
    return (plus - minus) * 0.5f;
}
__device__ __forceinline__ void quad_coords(uint32_t base, int w, int h, int& x0, int& y, int& z) {
    const uint32_t r = base / (uint32_t)w;
    x0 = (int)(base - r * (uint32_t)w);
    z = (int)(r / (uint32_t)h);
    y = (int)(r - (uint32_t)z * (uint32_t)h);
}

__global__ __launch_bounds__(kT) void k_stencil_g4(float* __restrict__ Gx, float* __restrict__ Gy,
                                                   float* __restrict__ Gz, FwdStencil fwd,
                                                   const float* __restrict__ phi, const float* __restrict__ gout,
                                                   int w, int h, int l, float dt, float nu, float mu, float eps,
                                                   const BlockDesc* __restrict__ bt) {
#pragma clang fp contract(off)
    if (bt) {   // blockIdx.y: the block, as in k_q_convx
        const int64_t o = bt[blockIdx.y].off;
        Gx += o;
        Gy += o;
        Gz += o;
        for (int a = 0; a < 3; ++a) {
            fwd.n[a] += o;
            if (fwd.g[a]) fwd.g[a] += o;
        }
        phi += o;
        gout += o;
    }
    const uint32_t W = (uint32_t)w, plane = W * (uint32_t)h, n = plane * (uint32_t)l;
    auto kbar = [&](float p, float u) { return (nu * dt * dirac(p, eps) - mu * dt) * u; };
    auto lbar = [&](float u) { return mu * dt * u; };
    for (uint32_t k = blockIdx.x * kT + threadIdx.x; k < n / 4; k += gridDim.x * kT) {
        const uint32_t base = 4 * k;
        int x0, y, z;
        quad_coords(base, w, h, x0, y, z);
        const uint32_t iym = y > 0 ? base - W : base, iyp = y + 1 < h ? base + W : base;
        const uint32_t izm = z > 0 ? base - plane : base, izp = z + 1 < l ? base + plane : base;
        float pr[6], ur[6], pym[4], pyp[4], pzm[4], pzp[4], uym[4], uyp[4], uzm[4], uzp[4];
        ld6(pr, phi, base, n);
        ld6(ur, gout, base, n);
        ld4(pym, phi, iym);
        ld4(pyp, phi, iyp);
        ld4(pzm, phi, izm);
        ld4(pzp, phi, izp);
        ld4(uym, gout, iym);
        ld4(uyp, gout, iyp);
        ld4(uzm, gout, izm);
        ld4(uzp, gout, izp);
        float kr[6], lr[6];
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            kr[t] = kbar(pr[t], ur[t]);
            lr[t] = lbar(ur[t]);
        }
        float oGx[4], oGy[4], oGz[4], n0[4], n1[4], n2[4], g0[4], g1[4], g2[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t j = base + i;
            const int x = x0 + i;
            const float nbx = dTx_v(kr[i], kr[i + 1], kr[i + 2], j, x, w, n);
            const float nby = dTyz_v(kbar(pym[i], uym[i]), kr[i + 1], kbar(pyp[i], uyp[i]), y, h);
            const float nbz = dTyz_v(kbar(pzm[i], uzm[i]), kr[i + 1], kbar(pzp[i], uzp[i]), z, l);
            const float lbx = dTx_v(lr[i], lr[i + 1], lr[i + 2], j, x, w, n);
            const float lby = dTyz_v(lbar(uym[i]), lr[i + 1], lbar(uyp[i]), y, h);
            const float lbz = dTyz_v(lbar(uzm[i]), lr[i + 1], lbar(uzp[i]), z, l);
            // grad_at: right x neighbour the voxel itself at x = w-1, left the previous
            // linear element; y and z from the clamped rows and planes
            Grad g;
            g.x = ((x == w - 1 ? pr[i + 1] : pr[i + 2]) - pr[i]) * 0.5f;
            g.y = (pyp[i] - pym[i]) * 0.5f;
            g.z = (pzp[i] - pzm[i]) * 0.5f;
            const float M = sqrtf(g.x * g.x + g.y * g.y + g.z * g.z);
            const float r = 1.0f / (M + kEta);
            const float c = M > 0.f ? (nbx * g.x + nby * g.y + nbz * g.z) * r * r / M : 0.f;
            oGx[i] = nbx * r - g.x * c + lbx;
            oGy[i] = nby * r - g.y * c + lby;
            oGz[i] = nbz * r - g.z * c + lbz;
            n0[i] = g.x / (M + kEta);
            n1[i] = g.y / (M + kEta);
            n2[i] = g.z / (M + kEta);
            g0[i] = g.x;
            g1[i] = g.y;
            g2[i] = g.z;
        }
        st4(Gx, base, oGx);
        st4(Gy, base, oGy);
        st4(Gz, base, oGz);
        st4(fwd.n[0], base, n0);
        st4(fwd.n[1], base, n1);
        st4(fwd.n[2], base, n2);
        if (fwd.g[0]) {   // the one-voxel finishing kernel reads them; the quad one forms them from phi
            st4(fwd.g[0], base, g0);
            st4(fwd.g[1], base, g1);
            st4(fwd.g[2], base, g2);
        }
    }
}

__global__ __launch_bounds__(kT) void k_finish4(float* __restrict__ gin, float* __restrict__ partials,
                                                const float* __restrict__ phi, const float* __restrict__ gout,
                                                const float* __restrict__ Gx, const float* __restrict__ Gy,
                                                const float* __restrict__ Gz, FwdStencil fwd,
                                                const float* __restrict__ Bh, const float* __restrict__ Ba,
                                                const void* __restrict__ img, int u8, const float* __restrict__ fo,
                                                const float* __restrict__ fi, int w, int h, int l, float dt, float nu,
                                                float mu, float eps, float l1, float l2, const BlockDesc* __restrict__ bt) {
#pragma clang fp contract(off)
    if (bt) {   // blockIdx.y: the block, as in k_q_convx; its partial rows follow the previous block's
        const int64_t o = bt[blockIdx.y].off;
        gin += o;
        phi += o;
        gout += o;
        Gx += o;
        Gy += o;
        Gz += o;
        for (int a = 0; a < 3; ++a) fwd.n[a] += o;
        Bh += o;
        Ba += o;
        img = static_cast<const char*>(img) + o * (u8 ? 1 : 4);
        fo += o;
        fi += o;
    }
    const uint32_t W = (uint32_t)w, plane = W * (uint32_t)h, n = plane * (uint32_t)l;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    for (uint32_t k = blockIdx.x * kT + threadIdx.x; k < n / 4; k += gridDim.x * kT) {
        const uint32_t base = 4 * k;
        int x0, y, z;
        quad_coords(base, w, h, x0, y, z);
        const uint32_t iym = y > 0 ? base - W : base, iyp = y + 1 < h ? base + W : base;
        const uint32_t izm = z > 0 ? base - plane : base, izp = z + 1 < l ? base + plane : base;
        // G along x, y, z; the forward normals: component x along the row, y at the
        // rows y+-1, z at the planes z+-1
        float gxr[6], n0r[6], gy0[4], gym[4], gyp[4], gz0[4], gzm[4], gzp[4];
        float n1m[4], n1p[4], n2m[4], n2p[4];
        ld6(gxr, Gx, base, n);
        ld6(n0r, fwd.n[0], base, n);
        ld4(gy0, Gy, base);
        ld4(gym, Gy, iym);
        ld4(gyp, Gy, iyp);
        ld4(gz0, Gz, base);
        ld4(gzm, Gz, izm);
        ld4(gzp, Gz, izp);
        ld4(n1m, fwd.n[1], iym);
        ld4(n1p, fwd.n[1], iyp);
        ld4(n2m, fwd.n[2], izm);
        ld4(n2p, fwd.n[2], izp);
        // the forward gradient's components at the Laplacian's neighbours, from phi
        // as grad_at forms them: along x at the linear neighbours (phi at base-2 ..
        // base+5, clamped to the block), along y and z from the rows and planes within
        // 2 (clamped: the one-sided rules at the faces)
        float fr[8], pym2[4], pyp2[4], pzm2[4], pzp2[4], pym1[4], pyp1[4], pzm1[4], pzp1[4];
        fr[0] = phi[base >= 2 ? base - 2 : 0];
        fr[1] = phi[base >= 1 ? base - 1 : 0];
        fr[6] = phi[base + 4 < n ? base + 4 : n - 1];
        fr[7] = phi[base + 5 < n ? base + 5 : n - 1];
        ld4(pym1, phi, iym);
        ld4(pyp1, phi, iyp);
        ld4(pzm1, phi, izm);
        ld4(pzp1, phi, izp);
        ld4(pym2, phi, y > 1 ? base - 2 * W : (y > 0 ? base - W : base));
        ld4(pyp2, phi, y + 2 < h ? base + 2 * W : (y + 1 < h ? base + W : base));
        ld4(pzm2, phi, z > 1 ? base - 2 * plane : (z > 0 ? base - plane : base));
        ld4(pzp2, phi, z + 2 < l ? base + 2 * plane : (z + 1 < l ? base + plane : base));
        float I[4], fo4[4], fi4[4], p4[4], u4[4], bh[4], ba[4];
        if (u8) {
            const uint32_t v = *reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(img) + base);
#pragma unroll
            for (int i = 0; i < 4; ++i) I[i] = (float)((v >> (8 * i)) & 255u);
        } else {
            ld4(I, static_cast<const float*>(img), base);
        }
        ld4(fo4, fo, base);
        ld4(fi4, fi, base);
        ld4(p4, phi, base);
        ld4(u4, gout, base);
        ld4(bh, Bh, base);
        ld4(ba, Ba, base);
#pragma unroll
        for (int i = 0; i < 4; ++i) fr[i + 2] = p4[i];
        // g_x at the linear element base-1+s: its right neighbour is itself at x = w-1
        // (base-1 at x0 = 0 is the previous row's last voxel), its left one the
        // previous element
        float g0v[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const int xq = x0 - 1 + s;
            const bool self = xq == w - 1 || xq == -1;
            g0v[s] = ((self ? fr[s + 1] : fr[s + 2]) - fr[s]) * 0.5f;
        }
        float o[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t j = base + i;
            const int x = x0 + i;
            const float st = dTx_v(gxr[i], gxr[i + 1], gxr[i + 2], j, x, w, n) + dTyz_v(gym[i], gy0[i], gyp[i], y, h) +
                             dTyz_v(gzm[i], gz0[i], gzp[i], z, l);
            const float kap = (((x == w - 1 ? n0r[i + 1] : n0r[i + 2]) - n0r[i]) + (n1p[i] - n1m[i]) +
                               (n2p[i] - n2m[i])) * 0.5f;
            const float g0R = x == w - 1 ? g0v[i + 1] : g0v[i + 2], g0L = j == 0 ? g0v[1] : g0v[i];
            const float g1R = y == h - 1 ? (p4[i] - pym1[i]) * 0.5f : (pyp2[i] - p4[i]) * 0.5f;
            const float g1L = y == 0 ? (pyp1[i] - p4[i]) * 0.5f : (p4[i] - pym2[i]) * 0.5f;
            const float g2R = z == l - 1 ? (p4[i] - pzm1[i]) * 0.5f : (pzp2[i] - p4[i]) * 0.5f;
            const float g2L = z == 0 ? (pzp1[i] - p4[i]) * 0.5f : (p4[i] - pzm2[i]) * 0.5f;
            const float lap = ((g0R - g0L) + (g1R - g1L) + (g2R - g2L)) * 0.5f;
            const float d1 = I[i] - fo4[i], d2 = I[i] - fi4[i];
            const float Eo = d1 * d1, Ei = d2 * d2;
            const float F = l1 * Eo - l2 * Ei;
            const float p0 = p4[i], u = u4[i];
            const float d = dirac(p0, eps);
            const float Fb = -dt * d * u;
            o[i] = u + dirac_prime(p0, eps) * (-dt * F + nu * dt * kap) * u + d * (bh[i] + I[i] * ba[i]) + st;
            s0 += Eo * Fb;
            s1 -= Ei * Fb;
            s2 += dt * d * kap * u;
            s3 += dt * (lap - kap) * u;
        }
        st4(gin, base, o);
    }
    block_partials(partials, s0, s1, s2, s3);
}

// grad_params[q] += sum over the partial rows, in double.  One workgroup: thread
// t sums rows t, t + kR, t + 2 kR, ... and the kR sums meet in a fixed tree, so
// the order (and the result) is the same on every call.
constexpr int kR = 256;
__global__ __launch_bounds__(kR) void k_reduce_params(float* gparams, const float* partials, int rows) {
    __shared__ double red[4][kR];
    const int t = threadIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int r = t; r < rows; r += kR) {
        const float4 v = *reinterpret_cast<const float4*>(partials + 4 * (int64_t)r);
        s[0] += (double)v.x;
        s[1] += (double)v.y;
        s[2] += (double)v.z;
        s[3] += (double)v.w;
    }
    for (int q = 0; q < 4; ++q) red[q][t] = s[q];
    __syncthreads();
    for (int o = kR / 2; o > 0; o >>= 1) {
        if (t < o)
            for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + o];
        __syncthreads();
    }
    if (t < 4) gparams[t] += (float)red[t][0];
}

}  // namespace

int pointwise_q(float* QH, float* QA, const float* phi, const float* gout, const void* img, int u8, const float* fo,
                const float* fi, const float* PH, const float* PHc, int w, int h, int l, int m, const rsf_params* p,
                hipStream_t s) {
    const int64_t n = (int64_t)w * h * l;
    const QArgs a{phi, gout, fo, fi, PH, PHc, img, u8, m, p->dt, p->lambda1, p->lambda2, p->epsilon};
    hipLaunchKernelGGL(k_pointwise_q, grid_for(n), dim3(kT), 0, s, QH, QA, a, w, h, l);
    RSF_HIP(hipGetLastError());
    return RSF_OK;
}

int blur_transpose(float* out, const float* in, float* t1, float* t2, int w, int h, int l, const Derived& d,
                   hipStream_t s) {
    const int64_t n = (int64_t)w * h * l;
    hipLaunchKernelGGL(k_conv_same_zero, grid_for(n), dim3(kT), 0, s, t1, in, w, h, l, 0, d.taps, d.P, d.m);
    hipLaunchKernelGGL(k_conv_same_zero, grid_for(n), dim3(kT), 0, s, t2, (const float*)t1, w, h, l, 1, d.taps, d.P,
                       d.m);
    hipLaunchKernelGGL(k_conv_same_zero, grid_for(n), dim3(kT), 0, s, out, (const float*)t2, w, h, l, 2, d.taps, d.P,
                       d.m);
    RSF_HIP(hipGetLastError());
    return RSF_OK;
}

int q_blur_transpose2(float* Bh, float* Ba, float* QH, float* QA, float* t1, float* t2, const float* phi,
                      const float* gout, const void* img, int u8, const float* fo, const float* fi, const float* PH,
                      const float* PHc, int w, int h, int l, const Derived& d, const rsf_params* p, hipStream_t s,
                      const BlockDesc* bt, int nb) {
    const QArgs a{phi, gout, fo, fi, PH, PHc, img, u8, d.m, p->dt, p->lambda1, p->lambda2, p->epsilon};
    // rows of up to 2048 voxels fit the fused kernel's 64 KB of LDS
    if (w <= 2048) {
        switch (d.P) {
            case 7: return q_conv2_axes<7>(Bh, Ba, a, t1, t2, w, h, l, d.taps, s, bt, nb);
            case 21: return q_conv2_axes<21>(Bh, Ba, a, t1, t2, w, h, l, d.taps, s, bt, nb);
            case 35: return q_conv2_axes<35>(Bh, Ba, a, t1, t2, w, h, l, d.taps, s, bt, nb);
            default: break;
        }
    }
    if (bt) return set_error(RSF_ERR_UNSUPPORTED, "no batched co-state pass for %d taps", d.P);
    RSF_TRY(pointwise_q(QH, QA, phi, gout, img, u8, fo, fi, PH, PHc, w, h, l, d.m, p, s));
    return blur_transpose2(Bh, Ba, QH, QA, t1, t2, w, h, l, d, s);
}

int blur_transpose2(float* Bh, float* Ba, const float* QH, const float* QA, float* t1, float* t2, int w, int h,
                    int l, const Derived& d, hipStream_t s) {
    switch (d.P) {
        case 7: return conv2_axes<7>(Bh, Ba, QH, QA, t1, t2, w, h, l, d.taps, s);
        case 21: return conv2_axes<21>(Bh, Ba, QH, QA, t1, t2, w, h, l, d.taps, s);
        case 35: return conv2_axes<35>(Bh, Ba, QH, QA, t1, t2, w, h, l, d.taps, s);
        default:
            RSF_TRY(blur_transpose(Bh, QH, t1, t2, w, h, l, d, s));
            return blur_transpose(Ba, QA, t1, t2, w, h, l, d, s);
    }
}

// the quad kernels' conditions: rows a multiple of 4 long, under 2^31 voxels, every
// pointer 16-byte aligned
bool quads_fit(int w, int64_t n, std::initializer_list<const void*> ptrs) {
    if (w % 4 != 0 || n >= (int64_t(1) << 31)) return false;
    for (const void* q : ptrs)
        if (reinterpret_cast<uintptr_t>(q) % 16 != 0) return false;
    return true;
}

bool quads_ok(int w, int64_t n, const void* img, int u8, std::initializer_list<const void*> ptrs) {
    return reinterpret_cast<uintptr_t>(img) % (u8 ? 4 : 16) == 0 && quads_fit(w, n, ptrs);
}

bool batch_ok(int P, int w, int64_t n, const void* img, int u8, std::initializer_list<const void*> ptrs) {
    return (P == 7 || P == 21 || P == 35) && w <= 2048 && quads_ok(w, n, img, u8, ptrs);
}

int stencil_g(float* Gx, float* Gy, float* Gz, FwdStencil fwd, const float* phi, const float* gout, int w, int h,
              int l, const rsf_params* p, hipStream_t s, const BlockDesc* bt, int nb) {
    const int64_t n = (int64_t)w * h * l;
    if (bt || quads_fit(w, n, {Gx, Gy, Gz, fwd.n[0], fwd.n[1], fwd.n[2], fwd.g[0], fwd.g[1], fwd.g[2], phi, gout})) {
        const dim3 grid(grid_for(n / 4).x, (unsigned)nb);
        hipLaunchKernelGGL(k_stencil_g4, grid, dim3(kT), 0, s, Gx, Gy, Gz, fwd, phi, gout, w, h, l, p->dt, p->nu,
                           p->mu, p->epsilon, bt);
    } else {
        if (!fwd.g[0]) return set_error(RSF_ERR_INVALID_ARGUMENT, "the one-voxel stencil kernel stores the gradients");
        hipLaunchKernelGGL(k_stencil_g, grid_for(n), dim3(kT), 0, s, Gx, Gy, Gz, fwd, phi, gout, w, h, l, p->dt,
                           p->nu, p->mu, p->epsilon);
    }
    RSF_HIP(hipGetLastError());
    return RSF_OK;
}

// the finishing kernel runs at most 1024 workgroups per block, each thread
// accumulating several voxels before the reduction (4096 workgroups at 128^3
// saved 1 us in it and cost 37 us in the parameter reduction)
dim3 grid_finish(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + kT - 1) / kT, 1024))); }

int finish_rows(int64_t n) { return (int)grid_finish(n).x; }

int finish(float* gin, float* partials, const float* phi, const float* gout, const float* Gx, const float* Gy,
           const float* Gz, FwdStencil fwd, const float* Bh, const float* Ba, const void* img, int u8,
           const float* fo, const float* fi, int w, int h, int l, const rsf_params* p, hipStream_t s,
           const BlockDesc* bt, int nb) {
    const int64_t n = (int64_t)w * h * l;
    // the same grid either way (finish_rows partial rows); a uint8 image is read 4 bytes at a time
    const bool img_ok = reinterpret_cast<uintptr_t>(img) % (u8 ? 4 : 16) == 0;
    if (bt || (img_ok && quads_fit(w, n, {gin, phi, gout, Gx, Gy, Gz, fwd.n[0], fwd.n[1], fwd.n[2], Bh, Ba, fo,
                                          fi}))) {
        const dim3 grid(grid_finish(n).x, (unsigned)nb);
        hipLaunchKernelGGL(k_finish4, grid, dim3(kT), 0, s, gin, partials, phi, gout, Gx, Gy, Gz, fwd, Bh, Ba, img, u8,
                           fo, fi, w, h, l, p->dt, p->nu, p->mu, p->epsilon, p->lambda1, p->lambda2, bt);
    } else {
        if (!fwd.g[0]) return set_error(RSF_ERR_INVALID_ARGUMENT, "the one-voxel finishing kernel reads the gradients");
        hipLaunchKernelGGL(k_finish, grid_finish(n), dim3(kT), 0, s, gin, partials, phi, gout, Gx, Gy, Gz, fwd, Bh,
                           Ba, img, u8, fo, fi, w, h, l, p->dt, p->nu, p->mu, p->epsilon, p->lambda1, p->lambda2);
    }
    RSF_HIP(hipGetLastError());
    return RSF_OK;
}

int reduce_params(float* gparams, const float* partials, int rows, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_params, dim3(1), dim3(kR), 0, s, gparams, partials, rows);
    RSF_HIP(hipGetLastError());
    return RSF_OK;
}

}  // namespace adj
}  // namespace rsf
